// Host ordering and symbolic analysis for the supernodal LDL^T.
//
// Works on a symmetric matrix given by its lower CSC pattern over scalar indices plus a partition
// of the scalars into "nodes" (consecutive ranges: the optimized keys for the full Hessian, the
// camera blocks for a reduced camera system). Ordering, elimination tree and supernodes are built
// on the node graph; a node's scalars stay consecutive and in their original order.
#pragma once

#include <cstdint>
#include <vector>

namespace symhip {

enum class Ordering { kNatural = 0, kMinDegree = 1, kGiven = 2, kNestedDissection = 3 };

struct CholSymbolic {
  int32_t n = 0;        // scalar dimension
  int32_t n_nodes = 0;  // node count
  std::vector<int32_t> node_ptr;  // [n_nodes+1] scalar ranges of nodes (original order)
  // Scalar permutation, reference convention (sparse_cholesky_solver.tcc:12-24):
  // inv_perm[new] = old (the ordering functor's output), perm[old] = new.
  std::vector<int32_t> perm, inv_perm;
  std::vector<int32_t> node_order;  // elimination order of nodes (new -> old node)

  // Supernodes over the permuted scalar index space.
  int32_t n_snodes = 0;
  std::vector<int32_t> sn_col_begin;  // [n_snodes+1] first permuted scalar column of supernode
  std::vector<int32_t> sn_parent;     // supernodal elimination tree (-1 root)
  std::vector<int32_t> sn_level;      // height above leaves (0 = leaf)
  std::vector<int64_t> sn_panel_off;  // [n_snodes+1] panel offsets (doubles), column-major
  std::vector<int32_t> sn_rows_ptr;   // [n_snodes+1] into sn_rows
  std::vector<int32_t> sn_rows;       // permuted scalar rows of each panel (own cols first)
  int32_t n_levels = 0;
  std::vector<int32_t> level_ptr, level_snodes;  // supernodes grouped by level

  // Right-looking update targets: for supernode s and each below-row segment that lands in one
  // ancestor t, the ancestor id, the segment's first below-row index, and for every below-row from
  // that segment on, its row position inside t's panel.
  std::vector<int32_t> upd_ptr;       // [n_snodes+1] into upd_target/upd_seg_begin
  std::vector<int32_t> upd_target, upd_seg_begin, upd_rel_ptr;
  std::vector<int32_t> upd_rel;       // row positions in target panel
  std::vector<int32_t> upd_col_off;   // column offset of the segment's first row in target

  int64_t panel_size() const { return sn_panel_off.empty() ? 0 : sn_panel_off.back(); }
  int64_t l_nnz = 0;       // nnz of strictly lower L implied by the supernodes (incl. padding)
  double factor_flops = 0; // flops of the numeric factorization (sum over supernodes)
  double update_flops = 0; // of which the trailing updates of ancestors (below-rows^2 x cols)
  int32_t max_snode_cols = 0, max_snode_rows = 0;
};

// Minimum-degree ordering of a node graph given as adjacency lists (symmetric, no self loops);
// node weights are the scalar dims. Returns the elimination order (new -> old node).
std::vector<int32_t> min_degree_order(const std::vector<std::vector<int32_t>>& adj,
                                      const std::vector<int32_t>& weight);

// Nested-dissection ordering of a node graph (the family METIS belongs to; the reference orders
// with METIS_NodeND through Eigen::MetisOrdering): recursive vertex separators taken from a BFS
// level structure rooted at a pseudo-peripheral node, parts first, separator last; subgraphs of
// at most leaf_nodes nodes, and subgraphs whose level separator outweighs the smaller part, are
// ordered by minimum degree. The shallower elimination tree gives the
// level-scheduled numeric factorization more independent supernodes per level.
// Maps a symhip_ordering code to the enum (throws on an unknown code).
Ordering ordering_from_code(int32_t code);

std::vector<int32_t> nested_dissection_order(const std::vector<std::vector<int32_t>>& adj,
                                             const std::vector<int32_t>& weight,
                                             int32_t leaf_nodes);

// Builds the symbolic factorization. `colptr/rowind` is the lower pattern (full diagonal
// present). node_ptr partitions 0..n into consecutive node ranges. When ordering == kGiven,
// `given_node_order` is the node elimination order. max_cols caps supernode width.
void chol_symbolic(int32_t n, const int32_t* colptr, const int32_t* rowind,
                   const std::vector<int32_t>& node_ptr, Ordering ordering,
                   const std::vector<int32_t>* given_node_order, int32_t max_cols,
                   CholSymbolic& out);

// For each lower-pattern entry q of the input matrix, its position in the panel storage.
std::vector<int64_t> chol_assembly_map(const CholSymbolic& s, const int32_t* colptr,
                                       const int32_t* rowind);

// Reference-convention elimination tree parent and column counts of the scalar factor for the
// permuted pattern (sparse_cholesky_solver.tcc:55-89), used to check the symbolic layer.
void scalar_etree_counts(int32_t n, const int32_t* colptr, const int32_t* rowind,
                         const std::vector<int32_t>& perm, std::vector<int32_t>& parent,
                         std::vector<int32_t>& counts);

// The reference's strictly lower L pattern (CSC, rows sorted) of P A P^T: what
// SparseCholeskySolver::L() holds after a factorization (sparse_cholesky_solver.tcc:26-219).
void scalar_factor_pattern(int32_t n, const int32_t* colptr, const int32_t* rowind,
                           const std::vector<int32_t>& perm, std::vector<int32_t>& l_colptr,
                           std::vector<int32_t>& l_rowind);

}  // namespace symhip
