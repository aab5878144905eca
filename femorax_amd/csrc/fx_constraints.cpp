// Constraint checks of the compile step: the models the reference refuses because OpenSees'
// TransformationConstraintHandler would silently compute something other than what the fix
// and equalDOF rows say (dof_numbering.cpp:44-262, 303-339).
//
// Two rules. (1) One equalDOF command per constrained node: OpenSees applies only the first
// MP constraint on a node, so a node tied by two commands is refused (:303-339). A command is
// a caller-given command id, or without ids a (retained, constrained) pair. (2) The
// numbering must agree with an emulation of the handler: nodes in ascending tag order (tag =
// model index + 1), each node's ID array built from its unconstrained DOFs and its MP's DOFs,
// fixes placed at the handler's position, the MP part copied from the retained node's ID
// array (doneID), and every nodal DOF read back through that layout (:61-150). Where the
// fixed set differs, or the handler's values and the equations are not one to one, the model
// is refused with the pattern that causes it (:162-237).
#include <algorithm>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "fx_internal.hpp"

namespace fx {

namespace {

constexpr int64_t kFixedDof = -1;

std::string dofs1(const std::vector<int32_t>& d0) {
  std::string s;
  for (size_t k = 0; k < d0.size(); ++k) s += (k ? " " : "") + std::to_string(d0[k] + 1);
  return s;
}

// Which refused pattern the model holds (the first found), as the reference words it.
std::string pattern(const fx_model_desc& m, int32_t node, int32_t dof) {
  std::map<int32_t, int32_t> retained_of;             // constrained tag -> retained tag
  std::map<int32_t, std::vector<int32_t>> dofs_of;    // constrained tag -> DOFs (0-based)
  for (int32_t r = 0; r < m.ntie; ++r) {
    retained_of[m.tie_constrained[r] + 1] = m.tie_retained[r] + 1;
    dofs_of[m.tie_constrained[r] + 1].push_back(m.tie_dof[r]);
  }
  auto is_fixed = [&](int32_t tag, int32_t d) {
    for (int32_t r = 0; r < m.nfix; ++r)
      if (m.fix_node[r] + 1 == tag && m.fix_dof[r] == d) return true;
    return false;
  };
  for (int32_t r = 0; r < m.nfix; ++r) {
    const int32_t c = m.fix_node[r] + 1, d = m.fix_dof[r];
    auto it = dofs_of.find(c);
    if (it == dofs_of.end() || std::find(it->second.begin(), it->second.end(), d) == it->second.end()) continue;
    const int32_t R = retained_of[c];
    if (is_fixed(R, d)) continue;
    return "node " + std::to_string(c) + " DOF " + std::to_string(d + 1) + " is fixed and also tied by equalDOF to node " +
           std::to_string(R) + " DOF " + std::to_string(d + 1) +
           ", which is free: in OpenSees the fix is silently lost. Fix the retained node's DOF instead, or remove "
           "that DOF from the tie";
  }
  for (const auto& [c, R] : retained_of) {
    auto rm = retained_of.find(R);
    if (rm == retained_of.end()) continue;
    const std::string chain = "node " + std::to_string(c) + " is tied to node " + std::to_string(R) + " (DOFs " +
                              dofs1(dofs_of[c]) + "), which is itself tied to node " + std::to_string(rm->second) +
                              " (DOFs " + dofs1(dofs_of[R]) + "). ";
    if (dofs_of[c] != dofs_of[R])
      return chain + "This is a partial chain: OpenSees connects the DOFs of node " + std::to_string(c) +
             " to the wrong DOFs. Tie it directly to the final retained node";
    if (R > c)
      return chain + "Node " + std::to_string(R) + " has a larger tag than node " + std::to_string(c) +
             ": OpenSees reads its equations before they are set and node " + std::to_string(c) +
             " is silently fixed. Tie it directly to the final retained node";
  }
  return "node " + std::to_string(node + 1) + " DOF " + std::to_string(dof + 1) +
         ": OpenSees' Transformation handler would give this DOF a different value than the fix and equalDOF "
         "constraints say. Tie every node directly to its final retained node and fix retained nodes";
}

}  // namespace

int check_equal_dof_commands(const fx_model_desc& m, std::string& err) {
  std::unordered_map<int32_t, std::vector<std::pair<int64_t, int32_t>>> cmds;  // constrained -> (id, retained)
  std::vector<int32_t> order;
  for (int32_t r = 0; r < m.ntie; ++r) {
    const int64_t id = m.tie_command ? static_cast<int64_t>(m.tie_command[r]) : -1 - static_cast<int64_t>(m.tie_retained[r]);
    auto& list = cmds[m.tie_constrained[r]];
    if (list.empty()) order.push_back(m.tie_constrained[r]);
    bool seen = false;
    for (const auto& c : list) seen |= c.first == id;
    if (!seen) list.emplace_back(id, m.tie_retained[r]);
  }
  for (int32_t c : order) {
    const auto& list = cmds[c];
    if (list.size() < 2) continue;
    std::string to;
    for (size_t k = 0; k < list.size(); ++k)
      to += (k ? (k + 1 == list.size() ? " and " : ", ") : "") + std::string("node ") + std::to_string(list[k].second + 1);
    err = "fx_compile_model: node " + std::to_string(c + 1) + " is constrained by " + std::to_string(list.size()) +
          " equalDOF commands (to " + to + "); OpenSees applies only the first. Tie the node with one command";
    return FX_EINVAL;
  }
  return FX_OK;
}

int check_transformation_handler(const fx_model_desc& m, const std::vector<int32_t>& node_rank,
                                 const std::vector<int32_t>& node_eq, std::string& err) {
  const int32_t N = m.nnode;
  // the MP of each constrained node: retained node (the last row's) and DOFs in row order
  std::vector<int32_t> retained(N, -1);
  std::vector<std::vector<int32_t>> mpd(N), sp(N);
  for (int32_t r = 0; r < m.ntie; ++r) {
    retained[m.tie_constrained[r]] = m.tie_retained[r];
    mpd[m.tie_constrained[r]].push_back(m.tie_dof[r]);
  }
  for (int32_t r = 0; r < m.nfix; ++r) sp[m.fix_node[r]].push_back(m.fix_dof[r]);
  // ID arrays: plain nodes by nodal DOF, transformed nodes by (unconstrained DOFs, MP DOFs)
  std::vector<std::vector<int64_t>> id(N);
  std::vector<std::vector<int32_t>> unc(N);
  for (int32_t i = 0; i < N; ++i) {
    if (retained[i] < 0) {
      id[i] = {3 * int64_t(i), 3 * int64_t(i) + 1, 3 * int64_t(i) + 2};
      for (int32_t s : sp[i]) id[i][s] = kFixedDof;
      continue;
    }
    const auto& D = mpd[i];
    for (int32_t d = 0; d < 3; ++d)
      if (std::find(D.begin(), D.end(), d) == D.end()) unc[i].push_back(d);
    const size_t nu = unc[i].size();
    id[i].assign(nu + D.size(), kFixedDof);
    for (size_t p = 0; p < nu; ++p) id[i][p] = 3 * int64_t(i) + unc[i][p];
    for (int32_t s : sp[i]) {
      int32_t loc = 0;  // the handler's position: unconstrained DOFs below s
      for (int32_t d = 0; d < s; ++d) loc += std::find(D.begin(), D.end(), d) == D.end();
      if (loc < static_cast<int32_t>(id[i].size())) id[i][loc] = kFixedDof;
    }
  }
  for (int32_t i = 0; i < N; ++i) {  // doneID, ascending tag
    if (retained[i] < 0) continue;
    const auto& other = id[retained[i]];
    const size_t nu = unc[i].size();
    for (size_t k = 0; k < mpd[i].size(); ++k) {
      const int32_t pos = mpd[i][k];
      id[i][nu + k] = pos < static_cast<int32_t>(other.size()) ? other[pos] : kFixedDof;
    }
  }
  std::unordered_map<int64_t, int32_t> eq_of;
  std::unordered_map<int32_t, int64_t> value_of;
  for (int32_t i = 0; i < N; ++i)
    for (int32_t d = 0; d < 3; ++d) {
      int64_t v;
      if (retained[i] < 0) {
        v = id[i][d];
      } else {
        const auto& D = mpd[i];
        auto k = std::find(D.begin(), D.end(), d);
        if (k == D.end())
          v = id[i][std::find(unc[i].begin(), unc[i].end(), d) - unc[i].begin()];
        else
          v = id[i][unc[i].size() + (k - D.begin())];
      }
      const int32_t e = node_eq[3 * static_cast<size_t>(node_rank[i]) + d];
      bool same = (v == kFixedDof) == (e < 0);
      if (same && e >= 0) {
        auto a = eq_of.emplace(v, e).first;
        auto b = value_of.emplace(e, v).first;
        same = a->second == e && b->second == v;
      }
      if (!same) {
        err = "fx_compile_model: " + pattern(m, i, d);
        return FX_EINVAL;
      }
    }
  return FX_OK;
}

}  // namespace fx
