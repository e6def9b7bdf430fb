// ct_treeio.hip -- host-side checks of an imported tree (no device code).
//
// A tree read from a file or handed over as arrays (the reference's own attributes,
// celltree.py:80-97) reaches the traversal kernels, which follow node links and cell ranges
// without bounds checks.  ct_check_tree establishes, in one pass over each array, that the
// arrays describe a tree the kernels can walk: every child link points forwards and inside
// the node array, every node but the root has exactly one parent, every node's cell range
// lies in [0, n), bb_indices is a permutation of [0, n), and every element index addresses a
// vertex (faces: -1 padding allowed past the first three entries).
#include <cstdint>
#include <vector>

#include "ct_internal.h"

extern "C" {

int ct_check_tree(const ct_node* nodes, int64_t n_nodes, const int32_t* bb_indices, const int32_t* elements,
                  int64_t n_elements, int32_t n_max_vert, int64_t n_vertices, int32_t allow_fill, int32_t* depth) {
  CT_REQUIRE(nodes && bb_indices && elements && depth, "null pointer");
  CT_REQUIRE(n_nodes >= 1 && n_nodes < ((int64_t)1 << 31), "n_nodes must be in [1, 2**31)");
  CT_REQUIRE(n_elements >= 1 && n_elements < ((int64_t)1 << 31), "n_elements must be in [1, 2**31)");
  CT_REQUIRE(n_max_vert >= 2 && n_max_vert <= CT_MAX_N_VERTEX, "n_max_vert must be in [2, %d]", CT_MAX_N_VERTEX);
  CT_REQUIRE(n_vertices >= 1 && n_vertices < ((int64_t)1 << 31), "n_vertices must be in [1, 2**31)");
  // Nodes, in index order: a child comes after its parent, so a node's depth is final
  // before its children are reached.
  std::vector<int32_t> level((size_t)n_nodes, -1);
  std::vector<uint8_t> parents((size_t)n_nodes, 0);
  level[0] = 0;
  int32_t deepest = 0;
  for (int64_t i = 0; i < n_nodes; ++i) {
    const ct_node& nd = nodes[i];
    if (nd.ptr < 0 || nd.size < 0 || (int64_t)nd.ptr + nd.size > n_elements) {
      ::ct::set_error("invalid tree: node %lld's cell range [%d, %d + %d) leaves [0, %lld)", (long long)i, nd.ptr,
                      nd.ptr, nd.size, (long long)n_elements);
      return CT_ERR_ARG;
    }
    if (nd.child == -1) continue;
    if (nd.child <= i || (int64_t)nd.child + 1 >= n_nodes) {
      ::ct::set_error("invalid tree: node %lld's child link %d points backwards or past the last node", (long long)i,
                      nd.child);
      return CT_ERR_ARG;
    }
    if (level[i] < 0) {
      ::ct::set_error("invalid tree: node %lld is not reached from the root", (long long)i);
      return CT_ERR_ARG;
    }
    for (int k = 0; k < 2; ++k) {
      const int64_t c = (int64_t)nd.child + k;
      if (parents[c]++) {
        ::ct::set_error("invalid tree: node %lld has two parents", (long long)c);
        return CT_ERR_ARG;
      }
      level[c] = level[i] + 1;
    }
    if (level[i] + 1 > deepest) deepest = level[i] + 1;
  }
  for (int64_t i = 1; i < n_nodes; ++i) {
    if (!parents[i]) {
      ::ct::set_error("invalid tree: node %lld is not reached from the root", (long long)i);
      return CT_ERR_ARG;
    }
  }
  // bb_indices: a permutation of [0, n).
  std::vector<uint8_t> seen((size_t)n_elements, 0);
  for (int64_t k = 0; k < n_elements; ++k) {
    const int32_t c = bb_indices[k];
    if (c < 0 || c >= n_elements || seen[c]++) {
      ::ct::set_error("bb_indices is not a permutation of range(%lld) (entry %lld = %d)", (long long)n_elements,
                      (long long)k, c);
      return CT_ERR_ARG;
    }
  }
  // Elements: vertex indices in range; faces may pad with -1 past their first three entries.
  const int lo = allow_fill ? -1 : 0;
  const int first = allow_fill ? (n_max_vert < 3 ? n_max_vert : 3) : n_max_vert;
  for (int64_t r = 0; r < n_elements; ++r) {
    const int32_t* row = elements + r * n_max_vert;
    for (int j = 0; j < n_max_vert; ++j) {
      const int32_t v = row[j];
      if (v >= n_vertices || v < (j < first ? 0 : lo)) {
        ::ct::set_error("element %lld entry %d = %d is out of range for %lld vertices", (long long)r, j, v,
                        (long long)n_vertices);
        return CT_ERR_ARG;
      }
    }
  }
  *depth = deepest;
  return CT_OK;
}

}  // extern "C"
