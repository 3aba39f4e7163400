// forest.hpp -- host-side forests of refinement trees over the n_intervals^dim
// box: the meshes of several levels of adaptive refinement (adapt_mesh,
// source/explicit_wave.cc:265-330, levels n_refinements .. n_refinements +
// n_adaptive_refinements, :272-273).
//
// Encoding (the C-ABI's): the base cells in lexicographic order (xi_0
// fastest), each tree depth first -- 1 = refined, its 2^dim children follow
// in lexicographic order; 0 = an active cell.  Active cells are numbered in
// that order; a cell's level is its depth.  Every node of the trees (active
// or refined) is kept in a hash map keyed by its level and its integer
// coordinates at that level, so that neighbour and balance queries are
// lookups.
#pragma once
#include <stdint.h>

#include <unordered_map>
#include <vector>

namespace exw {

struct ForestCell {
  int level;
  int64_t I[3];
};

class Forest {
 public:
  static constexpr int COORD_BITS = 19;  // n_intervals << level < 2^19
  static constexpr int MAX_LEVEL = 20;

  int dim = 0, n_intervals = 0;
  std::vector<ForestCell> cells;  // active cells in the encoding's order

  static uint64_t key(int l, const int64_t *I) {
    uint64_t k = (uint64_t)l;
    for (int d = 0; d < 3; ++d) k = (k << COORD_BITS) | (uint64_t)I[d];
    return k;
  }
  // -2: no such node; -1: a refined node; >= 0: the active cell's index
  int find(int l, const int64_t *I) const {
    for (int d = 0; d < dim; ++d)
      if (I[d] < 0 || I[d] >= ((int64_t)n_intervals << l)) return -2;
    auto it = nodes_.find(key(l, I));
    return it == nodes_.end() ? -2 : it->second;
  }
  int64_t extent(int l) const { return (int64_t)n_intervals << l; }

  // parses n_tree entries (n_tree < 0: as many as the trees take); false on
  // a malformed encoding or coordinates beyond COORD_BITS
  bool parse(int dim_, int n_intervals_, const int *tree, int64_t n_tree, int64_t *used) {
    dim = dim_;
    n_intervals = n_intervals_;
    cells.clear();
    nodes_.clear();
    if ((dim != 2 && dim != 3) || n_intervals < 1 || n_intervals >= (1 << COORD_BITS))
      return false;
    int64_t pos = 0, nb = 1;
    for (int d = 0; d < dim; ++d) nb *= n_intervals;
    for (int64_t B = 0; B < nb; ++B) {
      int64_t I[3] = {0, 0, 0}, t = B;
      for (int d = 0; d < dim; ++d) {
        I[d] = t % n_intervals;
        t /= n_intervals;
      }
      if (!parse_node(tree, n_tree, pos, 0, I)) return false;
    }
    if (used) *used = pos;
    return true;
  }

  // a tree from a node map: 'refined' holds the refined nodes; every other
  // node reached is active
  static int64_t emit(int dim, int n_intervals,
                      const std::unordered_map<uint64_t, int> &refined_or_active, int *out,
                      int64_t cap) {
    int64_t pos = 0, nb = 1;
    for (int d = 0; d < dim; ++d) nb *= n_intervals;
    for (int64_t B = 0; B < nb; ++B) {
      int64_t I[3] = {0, 0, 0}, t = B;
      for (int d = 0; d < dim; ++d) {
        I[d] = t % n_intervals;
        t /= n_intervals;
      }
      pos = emit_node(dim, refined_or_active, 0, I, out, pos, cap);
      if (pos < 0) return -1;
    }
    return pos;
  }

  const std::unordered_map<uint64_t, int> &nodes() const { return nodes_; }

 private:
  std::unordered_map<uint64_t, int> nodes_;

  bool parse_node(const int *tree, int64_t n_tree, int64_t &pos, int l, const int64_t *I) {
    if ((n_tree >= 0 && pos >= n_tree) || l > MAX_LEVEL) return false;
    if (extent(l) >= ((int64_t)1 << COORD_BITS)) return false;
    const int t = tree[pos++];
    if (t == 0) {
      nodes_[key(l, I)] = (int)cells.size();
      ForestCell c{l, {I[0], I[1], I[2]}};
      cells.push_back(c);
      return true;
    }
    if (t != 1) return false;
    nodes_[key(l, I)] = -1;
    for (int ch = 0; ch < (1 << dim); ++ch) {
      int64_t J[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) J[d] = 2 * I[d] + ((ch >> d) & 1);
      if (!parse_node(tree, n_tree, pos, l + 1, J)) return false;
    }
    return true;
  }

  static int64_t emit_node(int dim, const std::unordered_map<uint64_t, int> &m, int l,
                           const int64_t *I, int *out, int64_t pos, int64_t cap) {
    auto it = m.find(key(l, I));
    if (it == m.end()) return -1;
    const int refined = it->second == -1;
    if (out) {
      if (pos >= cap) return -1;
      out[pos] = refined;
    }
    ++pos;
    if (!refined) return pos;
    for (int ch = 0; ch < (1 << dim); ++ch) {
      int64_t J[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) J[d] = 2 * I[d] + ((ch >> d) & 1);
      pos = emit_node(dim, m, l + 1, J, out, pos, cap);
      if (pos < 0) return -1;
    }
    return pos;
  }
};

}  // namespace exw
