// Host ordering + symbolic analysis for the supernodal LDL^T (see chol_symbolic.h).
//
// The reference computes a METIS ordering and a scalar elimination tree + column counts
// (opt/sparse_cholesky/sparse_cholesky_solver.tcc:12-106). Here the ordering is a minimum-degree
// ordering of the node graph (METIS is not available; see the design notes), and the symbolic
// factorization is carried out on nodes, then grouped into supernodes whose panels are dense
// column-major blocks so the numeric phase runs as dense tile kernels on the GPU. The scalar
// etree/column counts of the reference are reproduced by scalar_etree_counts() for checking.
#include "./chol_symbolic.h"

#include <algorithm>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <string>

#include "./common.h"

namespace symhip {

std::vector<int32_t> min_degree_order(const std::vector<std::vector<int32_t>>& adj_in,
                                      const std::vector<int32_t>& weight) {
  const int32_t n = (int32_t)adj_in.size();
  std::vector<std::vector<int32_t>> adj = adj_in;
  for (auto& a : adj) std::sort(a.begin(), a.end());
  std::vector<int64_t> deg(n, 0);
  auto compute_deg = [&](int32_t v) {
    int64_t d = 0;
    for (int32_t u : adj[v]) d += weight[u];
    return d;
  };
  using Item = std::pair<int64_t, int32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  for (int32_t v = 0; v < n; ++v) {
    deg[v] = compute_deg(v);
    pq.push({deg[v], v});
  }
  std::vector<char> done(n, 0);
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<int32_t> merged;
  while (!pq.empty()) {
    const auto [d, v] = pq.top();
    pq.pop();
    if (done[v] || d != deg[v]) continue;
    done[v] = 1;
    order.push_back(v);
    const std::vector<int32_t> nv = adj[v];  // live neighbours
    for (int32_t u : nv) {
      // adj[u] <- (adj[u] U nv) \ {u, v}
      merged.clear();
      merged.reserve(adj[u].size() + nv.size());
      std::set_union(adj[u].begin(), adj[u].end(), nv.begin(), nv.end(),
                     std::back_inserter(merged));
      merged.erase(std::remove_if(merged.begin(), merged.end(),
                                  [&](int32_t x) { return x == u || x == v; }),
                   merged.end());
      adj[u].swap(merged);
      deg[u] = compute_deg(u);
      pq.push({deg[u], u});
    }
    adj[v].clear();
    adj[v].shrink_to_fit();
  }
  return order;
}

namespace {

// Recursive state for nested_dissection_order.
struct Dissector {
  const std::vector<std::vector<int32_t>>& adj;
  const std::vector<int32_t>& weight;
  int32_t leaf;
  std::vector<int32_t> order, stamp, dist;
  int32_t epoch = 0;

  Dissector(const std::vector<std::vector<int32_t>>& a, const std::vector<int32_t>& w, int32_t l)
      : adj(a), weight(w), leaf(l), stamp(a.size(), -1), dist(a.size(), -1) {
    order.reserve(a.size());
  }

  // Marks verts as the current subgraph (stamp == epoch).
  void mark(const std::vector<int32_t>& verts) {
    ++epoch;
    for (int32_t v : verts) stamp[v] = epoch;
  }

  // BFS inside the marked subgraph; returns nodes in BFS order, dist filled for them.
  std::vector<int32_t> bfs(int32_t root) {
    std::vector<int32_t> q{root};
    dist[root] = 0;
    const int32_t seen = -epoch - 2;  // distinct per epoch, never a valid stamp
    int32_t* st = stamp.data();
    st[root] = seen;
    for (size_t h = 0; h < q.size(); ++h) {
      const int32_t v = q[h];
      for (int32_t u : adj[v])
        if (st[u] == epoch) {
          st[u] = seen;
          dist[u] = dist[v] + 1;
          q.push_back(u);
        }
    }
    for (int32_t v : q) st[v] = epoch;  // restore membership
    return q;
  }

  void leaf_order(const std::vector<int32_t>& verts) {
    // minimum degree on the induced subgraph, in local indices
    std::vector<int32_t> local(verts.size());
    std::vector<std::vector<int32_t>> sub(verts.size());
    std::vector<int32_t> w(verts.size());
    mark(verts);
    std::vector<int32_t>& pos = dist;  // reuse as local index map
    for (size_t i = 0; i < verts.size(); ++i) pos[verts[i]] = (int32_t)i;
    for (size_t i = 0; i < verts.size(); ++i) {
      w[i] = weight[verts[i]];
      for (int32_t u : adj[verts[i]])
        if (stamp[u] == epoch) sub[i].push_back(pos[u]);
    }
    for (int32_t i : min_degree_order(sub, w)) order.push_back(verts[i]);
  }

  void dissect(std::vector<int32_t> verts) {
    if ((int32_t)verts.size() <= leaf) {
      leaf_order(verts);
      return;
    }
    mark(verts);
    // connected components: dissect each on its own
    {
      std::vector<std::vector<int32_t>> comps;
      const int32_t my_epoch = epoch;
      for (int32_t v : verts) {
        if (stamp[v] != my_epoch) continue;  // already in a found component
        std::vector<int32_t> comp = bfs(v);
        for (int32_t u : comp) stamp[u] = -1;  // remove from the pool
        comps.push_back(std::move(comp));
      }
      if (comps.size() > 1) {
        for (auto& c : comps) dissect(std::move(c));
        return;
      }
      mark(verts);
    }
    // pseudo-peripheral root: BFS again from the farthest node while the eccentricity grows
    std::vector<int32_t> q = bfs(verts[0]);
    int32_t ecc = dist[q.back()];
    for (int it = 0; it < 4; ++it) {
      std::vector<int32_t> q2 = bfs(q.back());
      const int32_t e2 = dist[q2.back()];
      q.swap(q2);  // dist now belongs to q
      if (e2 <= ecc) break;
      ecc = e2;
    }
    // level weights
    const int32_t depth = dist[q.back()];
    if (depth < 2) {  // (near-)clique: nothing to separate
      leaf_order(verts);
      return;
    }
    std::vector<int64_t> lw(depth + 1, 0);
    int64_t total = 0;
    for (int32_t v : q) {
      lw[dist[v]] += weight[v];
      total += weight[v];
    }
    // separator level: among levels 1..depth-1, the smallest level whose removal keeps both sides
    // within [25%, 75%] of the weight; the median level if none does
    int64_t below = 0;
    int32_t best = -1, median = 1;
    int64_t best_w = INT64_MAX;
    for (int32_t l = 0; l <= depth; ++l) {
      if (l >= 1 && l <= depth - 1) {
        const int64_t above = total - below - lw[l];
        if (below <= total / 2) median = l;
        if (4 * below >= total && 4 * above >= total && lw[l] < best_w) {
          best_w = lw[l];
          best = l;
        }
      }
      below += lw[l];
    }
    if (best < 0) best = median;
    std::vector<int32_t> A, B, sep;
    int64_t wa = 0, wb = 0, ws = 0;
    for (int32_t v : q) {
      if (dist[v] < best) {
        A.push_back(v);
        wa += weight[v];
      } else if (dist[v] > best) {
        B.push_back(v);
        wb += weight[v];
      } else {
        sep.push_back(v);
        ws += weight[v];
      }
    }
    // A level separator heavier than the smaller part (e.g. the landmarks of a bipartite
    // camera/landmark graph) would be eliminated last and fill in densely: order this subgraph
    // by minimum degree instead.
    if (ws > std::min(wa, wb)) {
      leaf_order(verts);
      return;
    }
    dissect(std::move(A));
    dissect(std::move(B));
    leaf_order(sep);
  }
};

}  // namespace

Ordering ordering_from_code(int32_t code) {
  switch (code) {
    case 0: return Ordering::kNatural;
    case 1: return Ordering::kMinDegree;
    case 2: return Ordering::kNestedDissection;
  }
  throw Error("symhip: unknown ordering " + std::to_string(code));
}

std::vector<int32_t> nested_dissection_order(const std::vector<std::vector<int32_t>>& adj,
                                             const std::vector<int32_t>& weight,
                                             int32_t leaf_nodes) {
  const int32_t n = (int32_t)adj.size();
  Dissector d(adj, weight, std::max(1, leaf_nodes));
  std::vector<int32_t> all(n);
  std::iota(all.begin(), all.end(), 0);
  d.dissect(std::move(all));
  SYMHIP_CHECK((int32_t)d.order.size() == n, "internal: nested dissection lost nodes");
  return d.order;
}

void scalar_etree_counts(int32_t n, const int32_t* colptr, const int32_t* rowind,
                         const std::vector<int32_t>& perm, std::vector<int32_t>& parent,
                         std::vector<int32_t>& counts) {
  // Upper pattern of P A P^T by columns (row < col), then the reference's path-following.
  std::vector<std::vector<int32_t>> upper(n);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t q = colptr[c]; q < colptr[c + 1]; ++q) {
      const int32_t r = rowind[q];
      const int32_t pr = perm.empty() ? r : perm[r], pc = perm.empty() ? c : perm[c];
      if (pr == pc) continue;
      upper[std::max(pr, pc)].push_back(std::min(pr, pc));
    }
  parent.assign(n, -1);
  counts.assign(n, 0);
  std::vector<int32_t> visited(n, -1);
  for (int32_t k = 0; k < n; ++k) {
    visited[k] = k;
    for (int32_t i0 : upper[k]) {
      int32_t i = i0;
      while (visited[i] != k) {
        if (parent[i] == -1) parent[i] = k;
        counts[i]++;
        visited[i] = k;
        i = parent[i];
      }
    }
  }
}

void scalar_factor_pattern(int32_t n, const int32_t* colptr, const int32_t* rowind,
                           const std::vector<int32_t>& perm, std::vector<int32_t>& l_colptr,
                           std::vector<int32_t>& l_rowind) {
  std::vector<int32_t> parent, counts;
  scalar_etree_counts(n, colptr, rowind, perm, parent, counts);
  l_colptr.assign(n + 1, 0);
  for (int32_t j = 0; j < n; ++j) l_colptr[j + 1] = l_colptr[j] + counts[j];
  l_rowind.assign(l_colptr[n], 0);
  std::vector<std::vector<int32_t>> upper(n);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t q = colptr[c]; q < colptr[c + 1]; ++q) {
      const int32_t r = rowind[q];
      const int32_t pr = perm.empty() ? r : perm[r], pc = perm.empty() ? c : perm[c];
      if (pr != pc) upper[std::max(pr, pc)].push_back(std::min(pr, pc));
    }
  // row k of L is the union of the etree paths from the entries of column k of the upper
  // pattern; visiting k in increasing order appends each column's rows sorted
  std::vector<int32_t> visited(n, -1), fill(l_colptr.begin(), l_colptr.end() - 1);
  for (int32_t k = 0; k < n; ++k) {
    visited[k] = k;
    for (int32_t i0 : upper[k])
      for (int32_t i = i0; visited[i] != k; i = parent[i]) {
        l_rowind[fill[i]++] = k;
        visited[i] = k;
      }
  }
}

void chol_symbolic(int32_t n, const int32_t* colptr, const int32_t* rowind,
                   const std::vector<int32_t>& node_ptr, Ordering ordering,
                   const std::vector<int32_t>* given_node_order, int32_t max_cols,
                   CholSymbolic& s) {
  s = CholSymbolic();
  s.n = n;
  s.node_ptr = node_ptr;
  const int32_t nn = (int32_t)node_ptr.size() - 1;
  s.n_nodes = nn;
  SYMHIP_CHECK(nn >= 0 && node_ptr.back() == n, "node partition does not cover the matrix");
  std::vector<int32_t> node_of(n), dim(nn);
  for (int32_t v = 0; v < nn; ++v) {
    dim[v] = node_ptr[v + 1] - node_ptr[v];
    SYMHIP_CHECK(dim[v] > 0 && dim[v] <= max_cols, "node dimension out of range");
    for (int32_t i = node_ptr[v]; i < node_ptr[v + 1]; ++i) node_of[i] = v;
  }
  // Node graph
  std::vector<std::vector<int32_t>> adj(nn);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t q = colptr[c]; q < colptr[c + 1]; ++q) {
      const int32_t a = node_of[c], b = node_of[rowind[q]];
      if (a != b) {
        adj[a].push_back(b);
        adj[b].push_back(a);
      }
    }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  // Ordering
  if (ordering == Ordering::kNatural) {
    s.node_order.resize(nn);
    std::iota(s.node_order.begin(), s.node_order.end(), 0);
  } else if (ordering == Ordering::kGiven) {
    SYMHIP_CHECK(given_node_order && (int32_t)given_node_order->size() == nn,
                 "given node ordering has the wrong size");
    s.node_order = *given_node_order;
  } else if (ordering == Ordering::kNestedDissection) {
    s.node_order = nested_dissection_order(adj, dim, 24);
  } else {
    s.node_order = min_degree_order(adj, dim);
  }
  std::vector<int32_t> node_new(nn, -1);
  for (int32_t i = 0; i < nn; ++i) {
    SYMHIP_CHECK(s.node_order[i] >= 0 && s.node_order[i] < nn && node_new[s.node_order[i]] < 0,
                 "node ordering is not a permutation");
    node_new[s.node_order[i]] = i;
  }
  // Scalar permutation: nodes in elimination order, scalars within a node in original order.
  s.perm.assign(n, 0);
  s.inv_perm.assign(n, 0);
  std::vector<int32_t> new_first(nn + 1, 0);  // first permuted scalar of new node index
  for (int32_t i = 0; i < nn; ++i) new_first[i + 1] = new_first[i] + dim[s.node_order[i]];
  for (int32_t i = 0; i < nn; ++i) {
    const int32_t v = s.node_order[i];
    for (int32_t k = 0; k < dim[v]; ++k) {
      s.perm[node_ptr[v] + k] = new_first[i] + k;
      s.inv_perm[new_first[i] + k] = node_ptr[v] + k;
    }
  }
  // Node-level symbolic factorization in new node indices.
  std::vector<std::vector<int32_t>> st(nn), children(nn);
  std::vector<int32_t> parent(nn, -1);
  std::vector<int32_t> mark(nn, -1);
  for (int32_t j = 0; j < nn; ++j) {
    std::vector<int32_t>& sj = st[j];
    mark[j] = j;
    for (int32_t u : adj[s.node_order[j]]) {
      const int32_t ju = node_new[u];
      if (ju > j && mark[ju] != j) {
        mark[ju] = j;
        sj.push_back(ju);
      }
    }
    for (int32_t c : children[j])
      for (int32_t r : st[c])
        if (r > j && mark[r] != j) {
          mark[r] = j;
          sj.push_back(r);
        }
    std::sort(sj.begin(), sj.end());
    if (!sj.empty()) {
      parent[j] = sj[0];
      children[sj[0]].push_back(j);
    }
  }
  // Supernodes: merge j-1 into j when j is its parent, width stays under max_cols, and the
  // explicit zeros added to j-1's columns stay small (relaxed amalgamation).
  std::vector<int32_t> sn_first_node;  // new node index of first node of each supernode
  int32_t cur_cols = 0;
  for (int32_t j = 0; j < nn; ++j) {
    const int32_t dj = dim[s.node_order[j]];
    bool merge = false;
    if (j > 0 && parent[j - 1] == j && cur_cols + dj <= max_cols) {
      // rows of j (below) vs rows of j-1 below j
      int64_t extra = 0;
      {
        // nodes in st[j] not in st[j-1]
        auto& a = st[j];
        auto& b = st[j - 1];
        size_t ib = 0;
        for (int32_t r : a) {
          while (ib < b.size() && b[ib] < r) ++ib;
          if (ib >= b.size() || b[ib] != r) extra += dim[s.node_order[r]];
        }
      }
      int64_t rows_j = dj;
      for (int32_t r : st[j]) rows_j += dim[s.node_order[r]];
      const int64_t zeros = extra * cur_cols;
      const int64_t total = rows_j * (cur_cols + dj);
      merge = (extra == 0) || (zeros * 4 <= total) || (zeros <= 256);
    }
    if (merge) {
      cur_cols += dj;
    } else {
      sn_first_node.push_back(j);
      cur_cols = dj;
    }
  }
  const int32_t ns = (int32_t)sn_first_node.size();
  s.n_snodes = ns;
  sn_first_node.push_back(nn);
  std::vector<int32_t> sn_of_node(nn);
  for (int32_t k = 0; k < ns; ++k)
    for (int32_t j = sn_first_node[k]; j < sn_first_node[k + 1]; ++j) sn_of_node[j] = k;
  s.sn_col_begin.resize(ns + 1);
  for (int32_t k = 0; k <= ns; ++k) s.sn_col_begin[k] = new_first[sn_first_node[k]];
  s.sn_parent.assign(ns, -1);
  s.sn_rows_ptr.assign(ns + 1, 0);
  s.sn_panel_off.assign(ns + 1, 0);
  std::vector<std::vector<int32_t>> below_nodes(ns);
  for (int32_t k = 0; k < ns; ++k) {
    const int32_t last = sn_first_node[k + 1] - 1;
    std::vector<int32_t> rows;
    for (int32_t j = sn_first_node[k]; j <= last; ++j)
      for (int32_t r : st[j])
        if (r > last) rows.push_back(r);
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    below_nodes[k] = rows;
    if (!rows.empty()) s.sn_parent[k] = sn_of_node[rows[0]];
    const int32_t ncols = s.sn_col_begin[k + 1] - s.sn_col_begin[k];
    int32_t nrows = ncols;
    for (int32_t r : rows) nrows += dim[s.node_order[r]];
    s.sn_rows_ptr[k + 1] = s.sn_rows_ptr[k] + nrows;
    s.sn_panel_off[k + 1] = s.sn_panel_off[k] + (int64_t)nrows * ncols;
    s.max_snode_cols = std::max(s.max_snode_cols, ncols);
    s.max_snode_rows = std::max(s.max_snode_rows, nrows);
    const double nc = ncols, mb = nrows - ncols;
    s.factor_flops += nc * nc * nc / 3.0 + mb * nc * nc + mb * mb * nc;
    s.update_flops += mb * mb * nc;
    s.l_nnz += (int64_t)nrows * ncols - (int64_t)ncols * (ncols + 1) / 2;
  }
  s.sn_rows.resize(s.sn_rows_ptr[ns]);
  for (int32_t k = 0; k < ns; ++k) {
    int32_t q = s.sn_rows_ptr[k];
    for (int32_t c = s.sn_col_begin[k]; c < s.sn_col_begin[k + 1]; ++c) s.sn_rows[q++] = c;
    for (int32_t r : below_nodes[k])
      for (int32_t x = new_first[r]; x < new_first[r + 1]; ++x) s.sn_rows[q++] = x;
  }
  // Levels
  s.sn_level.assign(ns, 0);
  for (int32_t k = 0; k < ns; ++k)  // children precede parents
    if (s.sn_parent[k] >= 0)
      s.sn_level[s.sn_parent[k]] = std::max(s.sn_level[s.sn_parent[k]], s.sn_level[k] + 1);
  s.n_levels = ns ? 1 + *std::max_element(s.sn_level.begin(), s.sn_level.end()) : 0;
  s.level_ptr.assign(s.n_levels + 1, 0);
  for (int32_t k = 0; k < ns; ++k) s.level_ptr[s.sn_level[k] + 1]++;
  for (int32_t l = 0; l < s.n_levels; ++l) s.level_ptr[l + 1] += s.level_ptr[l];
  s.level_snodes.resize(ns);
  {
    std::vector<int32_t> fill(s.level_ptr.begin(), s.level_ptr.end() - 1);
    for (int32_t k = 0; k < ns; ++k) s.level_snodes[fill[s.sn_level[k]]++] = k;
  }
  // Update targets
  auto row_pos = [&](int32_t t, int32_t row) -> int32_t {
    const int32_t c0 = s.sn_col_begin[t], c1 = s.sn_col_begin[t + 1];
    if (row >= c0 && row < c1) return row - c0;
    const int32_t* b = s.sn_rows.data() + s.sn_rows_ptr[t] + (c1 - c0);
    const int32_t* e = s.sn_rows.data() + s.sn_rows_ptr[t + 1];
    const int32_t* it = std::lower_bound(b, e, row);
    SYMHIP_CHECK(it != e && *it == row, "internal: update row missing from ancestor panel");
    return (int32_t)(it - (s.sn_rows.data() + s.sn_rows_ptr[t]));
  };
  s.upd_ptr.assign(ns + 1, 0);
  s.upd_rel_ptr.push_back(0);
  for (int32_t k = 0; k < ns; ++k) {
    const int32_t ncols = s.sn_col_begin[k + 1] - s.sn_col_begin[k];
    const int32_t* rows = s.sn_rows.data() + s.sn_rows_ptr[k] + ncols;
    const int32_t mb = s.sn_rows_ptr[k + 1] - s.sn_rows_ptr[k] - ncols;
    int32_t i = 0;
    while (i < mb) {
      const int32_t t = sn_of_node[node_new[node_of[s.inv_perm[rows[i]]]]];
      int32_t e = i;
      while (e < mb && rows[e] < s.sn_col_begin[t + 1]) ++e;
      s.upd_target.push_back(t);
      s.upd_seg_begin.push_back(i);
      s.upd_col_off.push_back(rows[i] - s.sn_col_begin[t]);
      for (int32_t x = i; x < mb; ++x) s.upd_rel.push_back(row_pos(t, rows[x]));
      s.upd_rel_ptr.push_back((int32_t)s.upd_rel.size());
      i = e;
    }
    s.upd_ptr[k + 1] = (int32_t)s.upd_target.size();
  }
}

std::vector<int64_t> chol_assembly_map(const CholSymbolic& s, const int32_t* colptr,
                                       const int32_t* rowind) {
  const int32_t n = s.n;
  // supernode of each permuted column
  std::vector<int32_t> sn_of_col(n);
  for (int32_t k = 0; k < s.n_snodes; ++k)
    for (int32_t c = s.sn_col_begin[k]; c < s.sn_col_begin[k + 1]; ++c) sn_of_col[c] = k;
  std::vector<int64_t> map(colptr[n]);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t q = colptr[c]; q < colptr[c + 1]; ++q) {
      int32_t pr = s.perm[rowind[q]], pc = s.perm[c];
      if (pr < pc) std::swap(pr, pc);  // lower triangle of the permuted matrix
      const int32_t t = sn_of_col[pc];
      const int32_t c0 = s.sn_col_begin[t], c1 = s.sn_col_begin[t + 1];
      int32_t rp;
      if (pr < c1) {
        rp = pr - c0;
      } else {
        const int32_t* b = s.sn_rows.data() + s.sn_rows_ptr[t] + (c1 - c0);
        const int32_t* e = s.sn_rows.data() + s.sn_rows_ptr[t + 1];
        const int32_t* it = std::lower_bound(b, e, pr);
        SYMHIP_CHECK(it != e && *it == pr, "internal: matrix entry outside the factor pattern");
        rp = (int32_t)(it - (s.sn_rows.data() + s.sn_rows_ptr[t]));
      }
      const int64_t ld = s.sn_rows_ptr[t + 1] - s.sn_rows_ptr[t];
      map[q] = s.sn_panel_off[t] + (int64_t)(pc - c0) * ld + rp;
    }
  return map;
}

}  // namespace symhip
