// Block decomposition for the deterministic LDS-staged scatter.
//
// The reference assembles element results through per-element slots and a per-equation gather
// in a fixed order (src/exec/assembly.cpp:28-38, pair order internal_force.cpp:121-134): every
// element value is written to and read back from HBM (24 doubles per element). Here the slot
// round trip stays in LDS: a workgroup owns a compact block of elements, writes their 24 values
// to LDS and sums them per node in a fixed order. A node only its block touches is written
// there; a node on a block boundary gets one partial-sum slot per block that touches it, and a
// second kernel adds a node's slots in block order. Every block runs in one launch, and the
// result is deterministic (the same bits for every run and co-runner) without atomics.
#include <algorithm>
#include <numeric>

#include "fx_internal.hpp"

namespace fx {

namespace {

struct Rcb {
  const std::vector<double>& cx;
  const std::vector<double>& cy;
  const std::vector<double>& cz;
  std::vector<std::vector<int32_t>>& out;

  void split(std::vector<int32_t>::iterator b, std::vector<int32_t>::iterator e) {
    const int64_t n = e - b;
    if (n <= kBlockElems) {
      std::vector<int32_t> blk(b, e);
      std::sort(blk.begin(), blk.end());
      out.push_back(std::move(blk));
      return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (auto it = b; it != e; ++it) {
      const double p[3] = {cx[*it], cy[*it], cz[*it]};
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], p[d]);
        hi[d] = std::max(hi[d], p[d]);
      }
    }
    int axis = 0;
    for (int d = 1; d < 3; ++d)
      if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
    const std::vector<double>& c = axis == 0 ? cx : axis == 1 ? cy : cz;
    const int64_t nb = (n + kBlockElems - 1) / kBlockElems;
    const int64_t left = (nb / 2) * kBlockElems;
    auto mid = b + left;
    std::nth_element(b, mid, e, [&](int32_t x, int32_t y) { return c[x] != c[y] ? c[x] < c[y] : x < y; });
    split(b, mid);
    split(mid, e);
  }
};

}  // namespace

void build_blocks(int32_t count, int32_t padded, const int32_t* node, int32_t nnode, const double* coords,
                  Blocks& B) {
  std::vector<double> cx(count), cy(count), cz(count);
  for (int32_t e = 0; e < count; ++e) {
    double s[3] = {0, 0, 0};
    for (int a = 0; a < 8; ++a) {
      const int32_t n = node[static_cast<size_t>(a) * padded + e];
      for (int d = 0; d < 3; ++d) s[d] += coords[static_cast<size_t>(d) * nnode + n];
    }
    cx[e] = s[0] / 8;
    cy[e] = s[1] / 8;
    cz[e] = s[2] / 8;
  }
  std::vector<int32_t> ids(count);
  std::iota(ids.begin(), ids.end(), 0);
  std::vector<std::vector<int32_t>> blocks;
  Rcb{cx, cy, cz, blocks}.split(ids.begin(), ids.end());
  const int32_t nb = static_cast<int32_t>(blocks.size());
  B.nblocks = nb;
  B.elems.assign(static_cast<size_t>(nb) * kBlockElems, -1);
  B.pos.assign(static_cast<size_t>(count), -1);
  B.slot_list.assign(static_cast<size_t>(nb) * 8 * kBlockElems, 0);
  B.lcorner.assign(static_cast<size_t>(nb) * 8 * kBlockElems, 0);
  B.ln_start.assign(static_cast<size_t>(nb) + 1, 0);
  B.lnode.clear();
  B.ln_slots.clear();
  std::vector<std::vector<int32_t>> block_nodes(nb);
  for (int32_t b = 0; b < nb; ++b) {
    const auto& el = blocks[b];
    std::vector<int32_t> nodes;
    nodes.reserve(el.size() * 8);
    for (size_t i = 0; i < el.size(); ++i) {
      B.elems[static_cast<size_t>(b) * kBlockElems + i] = el[i];
      B.pos[el[i]] = b * kBlockElems + static_cast<int32_t>(i);
      for (int a = 0; a < 8; ++a) nodes.push_back(node[static_cast<size_t>(a) * padded + el[i]]);
    }
    std::sort(nodes.begin(), nodes.end());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
    B.max_ln = std::max<int32_t>(B.max_ln, static_cast<int32_t>(nodes.size()));
    // local corner map and the per-local-node slot lists in (element, corner) order
    std::vector<std::vector<uint16_t>> per(nodes.size());
    for (size_t i = 0; i < el.size(); ++i)
      for (int a = 0; a < 8; ++a) {
        const int32_t n = node[static_cast<size_t>(a) * padded + el[i]];
        const auto ln = static_cast<uint16_t>(std::lower_bound(nodes.begin(), nodes.end(), n) - nodes.begin());
        per[ln].push_back(static_cast<uint16_t>(i * 8 + a));
        B.lcorner[(static_cast<size_t>(b) * 8 + a) * kBlockElems + i] = ln;
      }
    uint32_t pos = 0;
    for (size_t ln = 0; ln < nodes.size(); ++ln) {
      B.ln_slots.push_back(pos | (static_cast<uint32_t>(per[ln].size()) << 16));
      for (uint16_t s : per[ln]) B.slot_list[static_cast<size_t>(b) * 8 * kBlockElems + pos++] = s;
    }
    B.ln_start[b + 1] = B.ln_start[b] + static_cast<int32_t>(nodes.size());
    block_nodes[b] = std::move(nodes);
  }
  // nodes several blocks touch: one partial-sum slot per (node, block), a node's slots
  // consecutive and in block order
  std::vector<int32_t> nblk(nnode, 0);
  for (int32_t b = 0; b < nb; ++b)
    for (int32_t n : block_nodes[b]) ++nblk[n];
  std::vector<int32_t> first_slot(nnode, -1);
  B.sh_node.clear();
  B.sh_start.assign(1, 0);
  for (int32_t n = 0; n < nnode; ++n)
    if (nblk[n] > 1) {
      first_slot[n] = B.sh_start.back();
      B.sh_node.push_back(n);
      B.sh_start.push_back(B.sh_start.back() + nblk[n]);
    }
  std::vector<int32_t> taken(nnode, 0);
  B.lnode.assign(static_cast<size_t>(B.ln_start[nb]), 0);
  B.ln_part.assign(static_cast<size_t>(B.ln_start[nb]), -1);
  for (int32_t b = 0; b < nb; ++b) {
    const auto& nodes = block_nodes[b];
    for (size_t ln = 0; ln < nodes.size(); ++ln) {
      const int32_t n = nodes[ln];
      const size_t at = static_cast<size_t>(B.ln_start[b]) + ln;
      B.lnode[at] = static_cast<uint32_t>(n);
      if (first_slot[n] >= 0) B.ln_part[at] = first_slot[n] + taken[n]++;
    }
  }
}

void build_contraction(int32_t nnode, int32_t neq, const int32_t* node_eq, Contraction& out) {
  std::vector<int32_t> count(static_cast<size_t>(neq) + 1, 0);
  for (int32_t a = 0; a < 3 * nnode; ++a)
    if (node_eq[a] >= 0) ++count[node_eq[a] + 1];
  out.tied = false;
  for (int32_t q = 0; q < neq; ++q)
    if (count[q + 1] > 1) out.tied = true;
  out.start.assign(static_cast<size_t>(neq) + 1, 0);
  for (int32_t q = 0; q < neq; ++q) out.start[q + 1] = out.start[q] + count[q + 1];
  out.list.assign(static_cast<size_t>(out.start[neq]), 0);
  std::vector<int32_t> fill(out.start.begin(), out.start.end() - 1);
  for (int32_t a = 0; a < 3 * nnode; ++a)
    if (node_eq[a] >= 0) out.list[fill[node_eq[a]]++] = a;
}

}  // namespace fx
