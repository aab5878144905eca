// Lattice detection for the structured fast path.
//
// Femora-built soil models are mostly regular grids of axis-aligned bricks, often graded (layers
// of different element heights). When every element of a batch is an axis-aligned box spanning
// one interval of a common set of planes along each axis, with its nodes in the OpenSees Brick
// order, the kernels need neither the connectivity nor the coordinates: a brick of lattice cells
// is a workgroup's block, its nodes are addressed by lattice position, and the element geometry
// is three numbers (one cell size for a uniform lattice, the cell's plane spacings for a
// rectilinear one). Per block the device then holds
// one 32-bit word per lattice node (first equation | free-DOF mask | shared flag) and a 16-bit
// parameter index per cell. Anything else (distorted or rotated elements, holes inside
// a tile are fine, ties that merge a node's DOFs with another node's) uses the general blocks.
#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>

#include "fx_internal.hpp"

namespace fx {

bool build_lattice(int32_t count, int32_t padded, const int32_t* node, const int32_t* param, int32_t nnode,
                   const double* coords, const int32_t* node_eq, int32_t neq, bool xcd_order, int tile_z,
                   bool allow_rect, Lattice& L) {
  if (count <= 0 || (tile_z != 4 && tile_z != 8 && tile_z != 16)) return false;
  const double* X = coords;
  const double* Y = coords + nnode;
  const double* Z = coords + 2 * static_cast<size_t>(nnode);
  auto nd = [&](int a, int32_t e) { return node[static_cast<size_t>(a) * padded + e]; };
  static const int OX[8] = {0, 1, 1, 0, 0, 1, 1, 0}, OY[8] = {0, 0, 1, 1, 0, 0, 1, 1}, OZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  const double* C3[3] = {X, Y, Z};
  const int* O3[3] = {OX, OY, OZ};
  std::vector<int32_t> ci(count), cj(count), ck(count);
  int32_t* cidx[3] = {ci.data(), cj.data(), ck.data()};
  int32_t ni = 0, nj = 0, nk = 0;
  double h0[3] = {0, 0, 0};
  std::vector<double> planes[3];
  // a uniform lattice: element 0 fixes the cell size; every element an axis-aligned box of that
  // size at a whole number of cells from the lowest corner
  auto uniform_lattice = [&]() {
    const int32_t n0 = nd(0, 0);
    h0[0] = X[nd(1, 0)] - X[n0];
    h0[1] = Y[nd(3, 0)] - Y[n0];
    h0[2] = Z[nd(4, 0)] - Z[n0];
    if (!(h0[0] > 0 && h0[1] > 0 && h0[2] > 0)) return false;
    const double tol = 1e-9 * std::min(h0[0], std::min(h0[1], h0[2]));
    double lo[3] = {1e300, 1e300, 1e300};
    for (int32_t e = 0; e < count; ++e)
      for (int ax = 0; ax < 3; ++ax) lo[ax] = std::min(lo[ax], C3[ax][nd(0, e)]);
    int32_t* n3[3] = {&ni, &nj, &nk};
    *n3[0] = *n3[1] = *n3[2] = 0;
    for (int32_t e = 0; e < count; ++e) {
      const int32_t b = nd(0, e);
      for (int a = 0; a < 8; ++a)
        for (int ax = 0; ax < 3; ++ax)
          if (std::fabs(C3[ax][nd(a, e)] - C3[ax][b] - O3[ax][a] * h0[ax]) > tol) return false;
      for (int ax = 0; ax < 3; ++ax) {
        const double f = (C3[ax][b] - lo[ax]) / h0[ax], r = std::nearbyint(f);
        if (std::fabs(f - r) > 1e-6 || r > 1e6) return false;
        cidx[ax][e] = static_cast<int32_t>(r);
        *n3[ax] = std::max(*n3[ax], cidx[ax][e] + 1);
      }
    }
    return true;
  };
  // a rectilinear lattice: the planes along each axis are the distinct node coordinates of the
  // batch (merged within 1e-9 of the smallest cell size); every element spans one interval along
  // each axis with its nodes in the OpenSees Brick order
  auto rect_lattice = [&]() {
    double hmin = 1e300;
    for (int32_t e = 0; e < count; ++e)
      for (int ax = 0; ax < 3; ++ax) {
        const double h = C3[ax][nd(ax == 0 ? 1 : ax == 1 ? 3 : 4, e)] - C3[ax][nd(0, e)];
        if (!(h > 0)) return false;
        hmin = std::min(hmin, h);
      }
    const double tol = 1e-9 * hmin;
    for (int ax = 0; ax < 3; ++ax) {
      std::vector<double> v;
      v.reserve(static_cast<size_t>(count) * 2);
      for (int32_t e = 0; e < count; ++e) {
        v.push_back(C3[ax][nd(0, e)]);
        v.push_back(C3[ax][nd(6, e)]);
      }
      std::sort(v.begin(), v.end());
      planes[ax].clear();
      for (double c : v)
        if (planes[ax].empty() || c - planes[ax].back() > tol) planes[ax].push_back(c);
      if (planes[ax].size() < 2 || planes[ax].size() > 1000001) return false;
    }
    auto plane_of = [&](int ax, double c) -> int32_t {
      const auto& P = planes[ax];
      auto it = std::lower_bound(P.begin(), P.end(), c - tol);
      if (it == P.end() || std::fabs(*it - c) > tol) return -1;
      return static_cast<int32_t>(it - P.begin());
    };
    for (int32_t e = 0; e < count; ++e)
      for (int ax = 0; ax < 3; ++ax) {
        const int32_t i0 = plane_of(ax, C3[ax][nd(0, e)]);
        if (i0 < 0) return false;
        for (int a = 1; a < 8; ++a)
          if (plane_of(ax, C3[ax][nd(a, e)]) != i0 + O3[ax][a]) return false;
        cidx[ax][e] = i0;
      }
    ni = static_cast<int32_t>(planes[0].size()) - 1;
    nj = static_cast<int32_t>(planes[1].size()) - 1;
    nk = static_cast<int32_t>(planes[2].size()) - 1;
    for (int ax = 0; ax < 3; ++ax) h0[ax] = planes[ax][1] - planes[ax][0];
    return true;
  };
  const bool uniform = uniform_lattice();
  if (!uniform && !(allow_rect && rect_lattice())) return false;
  for (int32_t e = 0; e < count; ++e)
    if (param[e] < 0 || param[e] >= 0xffff) return false;
  // ties: two nodes share either all their equations (equalDOF in every free DOF, the laminar
  // boundary of a soil column: the same equation triple) or none; tied[n]: another node holds n's
  // equations. Partial ties go to the general path.
  std::vector<char> tied(nnode, 0);
  {
    std::unordered_map<int32_t, int32_t> owner;
    for (int32_t n = 0; n < nnode; ++n)
      for (int d = 0; d < 3; ++d) {
        const int32_t q = node_eq[3 * n + d];
        if (q < 0) continue;
        auto it = owner.emplace(q, n);
        const int32_t m = it.first->second;
        if (it.second || m == n) continue;
        for (int c = 0; c < 3; ++c)
          if (node_eq[3 * m + c] != node_eq[3 * n + c]) return false;
        tied[m] = tied[n] = 1;
      }
  }
  // lattice -> element and -> node
  const int64_t ncell = static_cast<int64_t>(ni) * nj * nk;
  if (ncell > 4 * static_cast<int64_t>(count) + 4096) return false;  // too sparse to be a lattice
  std::vector<int32_t> cell(ncell, -1);
  const int32_t NI = ni + 1, NJ = nj + 1;
  std::vector<int32_t> lat_node(static_cast<size_t>(NI) * NJ * (nk + 1), -1);
  for (int32_t e = 0; e < count; ++e) {
    const int64_t c = ci[e] + static_cast<int64_t>(ni) * (cj[e] + static_cast<int64_t>(nj) * ck[e]);
    if (cell[c] >= 0) return false;  // two elements in one cell
    cell[c] = e;
    for (int a = 0; a < 8; ++a) {
      const size_t p = (ci[e] + OX[a]) + static_cast<size_t>(NI) * ((cj[e] + OY[a]) + static_cast<size_t>(NJ) * (ck[e] + OZ[a]));
      const int32_t n = nd(a, e);
      if (lat_node[p] >= 0 && lat_node[p] != n) return false;
      lat_node[p] = n;
    }
  }
  // tiles of kTile cells; tiles of one parity colour share no node, so the (at most 8) tiles at
  // a node have 8 different colours and their partial sums are added in colour order
  const int TX = kTileX, TY = kTileY, TZ = tile_z;
  const int tile_nodes_full = (TX + 1) * (TY + 1) * (TZ + 1);
  const int32_t bi = (ni + TX - 1) / TX, bj = (nj + TY - 1) / TY, bk = (nk + TZ - 1) / TZ;
  L = Lattice{};
  L.tz = tile_z;
  L.hx = h0[0];
  L.hy = h0[1];
  L.hz = h0[2];
  L.rect = !uniform;
  if (L.rect) {
    L.px = planes[0];
    L.py = planes[1];
    L.pz = planes[2];
    L.ecell.resize(static_cast<size_t>(count) * 3);
    for (int32_t e = 0; e < count; ++e) {
      L.ecell[3 * static_cast<size_t>(e)] = ci[e];
      L.ecell[3 * static_cast<size_t>(e) + 1] = cj[e];
      L.ecell[3 * static_cast<size_t>(e) + 2] = ck[e];
    }
  }
  std::vector<int32_t> order;
  order.reserve(static_cast<size_t>(bi) * bj * bk);
  if (xcd_order) {
    // Workgroups of a launch go round-robin to the 8 XCDs (each with its own L2): the tiles of
    // one of 8 slabs along the longest tile axis sit at positions r, r + 8, r + 16, ... so that
    // tiles sharing nodes meet in one L2. The tile order changes no sum (colour order does).
    const int axis = (bi >= bj && bi >= bk) ? 0 : (bj >= bk ? 1 : 2);
    const int32_t len = axis == 0 ? bi : axis == 1 ? bj : bk;
    std::vector<std::vector<int32_t>> slab(8);
    for (int32_t t = 0; t < bi * bj * bk; ++t) {
      const int32_t ti = t % bi, tj = (t / bi) % bj, tk = t / (bi * bj);
      const int32_t pos = axis == 0 ? ti : axis == 1 ? tj : tk;
      slab[std::min<int64_t>(7, static_cast<int64_t>(pos) * 8 / std::max<int32_t>(len, 1))].push_back(t);
    }
    std::vector<size_t> at(8, 0);
    for (bool more = true; more;) {
      more = false;
      for (int r = 0; r < 8; ++r)
        if (at[r] < slab[r].size()) {
          order.push_back(slab[r][at[r]++]);
          more = true;
        }
    }
  } else {
    for (int32_t t = 0; t < bi * bj * bk; ++t) order.push_back(t);
  }
  std::vector<int32_t> tile_nodes;  // global node (or -1) of every tile node, in word order
  std::vector<int8_t> tile_colour;
  for (int32_t t : order) {
    const int32_t ti = t % bi, tj = (t / bi) % bj, tk = t / (bi * bj);
    const int32_t i0 = ti * TX, j0 = tj * TY, k0 = tk * TZ;
    const int32_t nx = std::min(TX, ni - i0), ny = std::min(TY, nj - j0), nz = std::min(TZ, nk - k0);
    TileHeader h{};
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.ox = i0;
    h.oy = j0;
    h.oz = k0;
    // every tile owns a full block of 81 (TZ + 1) words (node (x, y, z) at x + 9 (y + 9 z), also
    // in a tile cut by the lattice's end: its missing nodes have no word) and 64 TZ cell
    // parameters in full-tile order (x + 8 (y + 8 z)), so a workgroup finds its words and cells
    // from the tile index alone and addresses its nodes with constant strides
    h.word_offset = static_cast<int32_t>(tile_nodes.size());
    h.cell_offset = static_cast<int32_t>(L.cell_param.size());
    bool any = false;
    for (int32_t z = 0; z < TZ; ++z)
      for (int32_t y = 0; y < TY; ++y)
        for (int32_t x = 0; x < TX; ++x) {
          int32_t e = -1;
          if (x < nx && y < ny && z < nz)
            e = cell[(i0 + x) + static_cast<int64_t>(ni) * ((j0 + y) + static_cast<int64_t>(nj) * (k0 + z))];
          L.cell_param.push_back(e >= 0 ? static_cast<uint16_t>(param[e]) : uint16_t(0xffff));
          L.cell_elem.push_back(e);
          any |= e >= 0;
        }
    if (!any) {
      L.cell_param.resize(h.cell_offset);
      L.cell_elem.resize(h.cell_offset);
      continue;
    }
    tile_nodes.resize(static_cast<size_t>(h.word_offset) + tile_nodes_full, -1);
    for (int32_t z = 0; z <= nz; ++z)
      for (int32_t y = 0; y <= ny; ++y)
        for (int32_t x = 0; x <= nx; ++x)
          tile_nodes[static_cast<size_t>(h.word_offset) + x + (TX + 1) * (y + (TY + 1) * z)] =
              lat_node[(i0 + x) + static_cast<size_t>(NI) * ((j0 + y) + static_cast<size_t>(NJ) * (k0 + z))];
    L.tiles.push_back(h);
    tile_colour.push_back(static_cast<int8_t>((ti & 1) | ((tj & 1) << 1) | ((tk & 1) << 2)));
  }
  // how many tiles touch each node, and which (in parity-colour order: a node's tiles all differ
  // in colour)
  const int32_t nt = static_cast<int32_t>(L.tiles.size());
  std::vector<int32_t> touch_count(nnode, 0);
  for (int32_t n : tile_nodes)
    if (n >= 0) ++touch_count[n];
  std::vector<int32_t> touch_start(static_cast<size_t>(nnode) + 1, 0);
  // a node gets partial-sum slots where several tiles touch it or it is tied
  auto slotted = [&](int32_t n) { return touch_count[n] > 1 || tied[n]; };
  for (int32_t n = 0; n < nnode; ++n) touch_start[n + 1] = touch_start[n] + (slotted(n) ? touch_count[n] : 0);
  std::vector<int32_t> touch_tile(touch_start[nnode]), touch_ln(touch_start[nnode]);
  {
    std::vector<int32_t> fill(touch_start.begin(), touch_start.end() - 1);
    for (int c = 0; c < 8; ++c)
      for (int32_t t = 0; t < nt; ++t) {
        if (tile_colour[t] != c) continue;
        const TileHeader& h = L.tiles[t];
        for (int32_t i = 0; i < tile_nodes_full; ++i) {
          const int32_t n = tile_nodes[static_cast<size_t>(h.word_offset) + i];
          if (n < 0 || !slotted(n)) continue;
          touch_tile[fill[n]] = t;
          touch_ln[fill[n]++] = i;
        }
      }
  }
  // shared nodes grouped by their set of tiles (a face, an edge or a corner between tiles), groups
  // and their nodes in first-touch order. A group of k tiles and m nodes keeps one block of m
  // partial-sum slots per tile: slot (base + j m + i) holds tile j's sum at node i.
  std::vector<int32_t> shared_nodes;
  {
    std::vector<char> seen(nnode, 0);
    std::map<std::vector<int32_t>, int32_t> group_of;
    std::vector<std::vector<int32_t>> groups;
    for (int32_t n : tile_nodes) {
      if (n < 0 || touch_count[n] < 2 || tied[n] || seen[n]) continue;
      seen[n] = 1;
      std::vector<int32_t> key(touch_tile.begin() + touch_start[n], touch_tile.begin() + touch_start[n + 1]);
      auto it = group_of.find(key);
      if (it == group_of.end()) {
        it = group_of.emplace(std::move(key), static_cast<int32_t>(groups.size())).first;
        groups.emplace_back();
      }
      groups[it->second].push_back(n);
    }
    for (const auto& g : groups) shared_nodes.insert(shared_nodes.end(), g.begin(), g.end());
  }
  // tied nodes grouped by their equations (one set per tied equation triple), sets and their
  // nodes in first-touch order
  std::vector<std::vector<int32_t>> tied_sets;
  {
    std::unordered_map<int32_t, int32_t> set_of;  // first free equation -> set
    std::vector<char> seen(nnode, 0);
    for (int32_t n : tile_nodes) {
      if (n < 0 || !tied[n] || seen[n]) continue;
      seen[n] = 1;
      int32_t key = -1;
      for (int d = 0; d < 3 && key < 0; ++d) key = node_eq[3 * n + d];
      auto it = set_of.emplace(key, static_cast<int32_t>(tied_sets.size()));
      if (it.second) tied_sets.emplace_back();
      tied_sets[it.first->second].push_back(n);
    }
  }
  // internal numbering: first the nodes only one tile touches (tile by tile, lattice order),
  // then the shared nodes in group order, then the tied sets, then untouched equations. A node's
  // free DOFs are consecutive (a tied set's nodes hold the same ones). perm: internal ->
  // reference equation.
  std::vector<int32_t> internal_of(static_cast<size_t>(neq), -1);
  L.perm.assign(static_cast<size_t>(neq), -1);
  int32_t next = 0;
  std::vector<char> numbered(nnode, 0);
  auto number = [&](int32_t n) {
    numbered[n] = 1;
    for (int d = 0; d < 3; ++d) {
      const int32_t q = node_eq[3 * n + d];
      if (q < 0) continue;
      internal_of[q] = next;
      L.perm[next++] = q;
    }
  };
  // a tile's nodes only it touches, tile by tile: its strictly-interior nodes first (x fastest),
  // then the others (on the lattice's boundary); a full tile whose interior nodes all have three
  // free DOFs then finds their equations from one base (TileHeader::interior_base)
  const int NXY = (TX + 1) * (TY + 1);
  for (TileHeader& h : L.tiles) {
    h.interior_base = -1;
    const bool full = h.nx == TX && h.ny == TY && h.nz == TZ;
    bool regular = full;
    const int32_t base = next;
    for (int32_t z = 1; z < TZ; ++z)
      for (int32_t y = 1; y < TY; ++y)
        for (int32_t x = 1; x < TX; ++x) {
          const int32_t n = tile_nodes[static_cast<size_t>(h.word_offset) + x + (TX + 1) * y + NXY * z];
          if (n < 0 || numbered[n] || touch_count[n] != 1 || tied[n]) {
            regular = false;
            continue;
          }
          bool all = true;
          for (int d = 0; d < 3; ++d) all &= node_eq[3 * n + d] >= 0;
          regular &= all;
          number(n);
        }
    if (regular && next == base + 3 * (TX - 1) * (TY - 1) * (TZ - 1)) h.interior_base = base;
    for (int32_t i = 0; i < tile_nodes_full; ++i) {
      const int32_t n = tile_nodes[static_cast<size_t>(h.word_offset) + i];
      if (n >= 0 && !numbered[n] && touch_count[n] == 1 && !tied[n]) number(n);
    }
  }
  for (int32_t n : shared_nodes) number(n);
  for (const auto& s : tied_sets) number(s.front());
  for (int32_t q = 0; q < neq; ++q)
    if (internal_of[q] < 0) {
      internal_of[q] = next;
      L.perm[next++] = q;
    }
  auto word_of = [&](int32_t n, uint32_t& w) {
    int32_t base = -1;
    uint32_t mask = 0;
    for (int d = 0; d < 3; ++d) {
      const int32_t q = node_eq[3 * n + d];
      if (q < 0) continue;
      if (base < 0) base = internal_of[q];
      mask |= 1u << d;
    }
    if (base >= (1 << 28)) return false;
    w = (mask << 28) | static_cast<uint32_t>(base < 0 ? 0 : base);
    return true;
  };
  // slots: group by group (nodes of a group are consecutive in shared_nodes)
  std::vector<int32_t> slot_of_touch(touch_tile.size(), -1);
  L.shared_start.clear();
  L.shared_meta.clear();
  L.shared_word.clear();
  int64_t nslot = 0;
  for (size_t g0 = 0; g0 < shared_nodes.size();) {
    const int32_t first = shared_nodes[g0];
    const int32_t k = touch_count[first];
    size_t g1 = g0 + 1;
    while (g1 < shared_nodes.size()) {
      const int32_t n = shared_nodes[g1];
      if (touch_count[n] != k ||
          !std::equal(touch_tile.begin() + touch_start[n], touch_tile.begin() + touch_start[n + 1],
                      touch_tile.begin() + touch_start[first]))
        break;
      ++g1;
    }
    const int32_t m = static_cast<int32_t>(g1 - g0);
    for (int32_t i = 0; i < m; ++i) {
      const int32_t n = shared_nodes[g0 + i];
      for (int32_t j = 0; j < k; ++j) slot_of_touch[touch_start[n] + j] = static_cast<int32_t>(nslot + static_cast<int64_t>(j) * m + i);
      uint32_t w = 0;
      if (!word_of(n, w)) return false;
      L.shared_word.push_back(w);
      L.shared_start.push_back(static_cast<int32_t>(nslot + i));
      L.shared_meta.push_back(static_cast<uint32_t>(m) | (static_cast<uint32_t>(k) << 24));
    }
    nslot += static_cast<int64_t>(k) * m;
    if (m >= (1 << 24) || nslot >= (int64_t(1) << 28)) return false;
    g0 = g1;
  }
  // tied sets: one slot per (node, touching tile), a set's slots contiguous (nodes in set order,
  // each node's tiles in colour order)
  L.tied_word.clear();
  L.tied_start.assign(1, static_cast<int32_t>(nslot));
  for (const auto& s : tied_sets) {
    uint32_t w = 0;
    if (!word_of(s.front(), w)) return false;
    L.tied_word.push_back(w);
    for (int32_t n : s)
      for (int32_t j = touch_start[n]; j < touch_start[n + 1]; ++j) slot_of_touch[j] = static_cast<int32_t>(nslot++);
    if (nslot >= (int64_t(1) << 28)) return false;
    L.tied_start.push_back(static_cast<int32_t>(nslot));
  }
  L.nslot = nslot;
  if (nslot >= (int64_t(1) << 28)) return false;  // a shared node's slot takes the word's 28 equation bits
  // tile node words: first equation | free-DOF mask | shared flag; and for the shared nodes the
  // tile's partial-sum slot
  L.slots.assign(tile_nodes.size(), -1);
  for (int32_t n = 0; n < nnode; ++n)
    for (int32_t j = touch_start[n]; j < touch_start[n + 1]; ++j)
      L.slots[static_cast<size_t>(L.tiles[touch_tile[j]].word_offset) + touch_ln[j]] = slot_of_touch[j];
  L.words.resize(tile_nodes.size());
  for (size_t i = 0; i < tile_nodes.size(); ++i) {
    const int32_t n = tile_nodes[i];
    uint32_t w = 0;
    if (n >= 0) {
      if (!word_of(n, w)) return false;
      if (slotted(n)) w |= 0x80000000u;
    }
    L.words[i] = w;
  }
  return true;
}

}  // namespace fx
