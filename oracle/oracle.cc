// Oracle (test infrastructure only) -- CPU FP64 restatement of the SymForce optimizer hot path.
// Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
//
// Restated files (reference root relative):
//   opt/factor.h:423-449            ComputeKeysToOptimize
//   opt/linearizer.cc:54-473        Relinearize / BuildInitialLinearization / dense scatter
//   opt/internal/linearizer_utils.h:82-252   storage maps, key helpers, factor offsets
//   opt/linearization.h:51-67       Error, LinearDeltaError
//   opt/sparse_cholesky/sparse_cholesky_solver.tcc:12-259   ordering, symbolic, LDL^T, solve
//   opt/sparse_schur_solver.tcc:16-163       Schur symbolic / factorize / solve
//   opt/levenberg_marquardt_solver.tcc:22-343, opt/internal/levenberg_marquardt_state.h:33-270
//   opt/internal/optimizer_utils.h:36-104, opt/optimizer.tcc:21-89, opt/optimizer.cc:8-56
//   opt/values.cc:294-327           Retract
#include "./oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "./factors.h"

namespace oracle {

static thread_local std::string g_error;

struct KeyT {
  int8_t letter;
  int64_t sub, super;
};
static bool lexical_less(const KeyT& a, const KeyT& b) {  // opt/key.cc:18-21
  return std::make_tuple(a.letter, a.sub, a.super) < std::make_tuple(b.letter, b.sub, b.super);
}

// ---------------------------------------------------------------------------------------------
// CSC helpers
// ---------------------------------------------------------------------------------------------
struct Csc {
  int32_t rows = 0, cols = 0;
  std::vector<int32_t> colptr, rowind;
  std::vector<double> val;
  int64_t nnz() const { return (int64_t)rowind.size(); }
};

// Eigen setFromTriplets: compressed, sorted by (col,row), duplicates summed in insertion order.
struct Triplet {
  int32_t r, c;
  double v;
};
static Csc csc_from_triplets(int32_t rows, int32_t cols, const std::vector<Triplet>& t) {
  std::vector<int64_t> idx(t.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    return std::make_pair(t[a].c, t[a].r) < std::make_pair(t[b].c, t[b].r);
  });
  Csc m;
  m.rows = rows;
  m.cols = cols;
  m.colptr.assign(cols + 1, 0);
  for (size_t k = 0; k < idx.size(); ++k) {
    const Triplet& e = t[idx[k]];
    if (!m.rowind.empty() && k > 0 && t[idx[k - 1]].c == e.c && t[idx[k - 1]].r == e.r) {
      m.val.back() += e.v;
      continue;
    }
    m.rowind.push_back(e.r);
    m.val.push_back(e.v);
    m.colptr[e.c + 1]++;
  }
  for (int32_t c = 0; c < cols; ++c) m.colptr[c + 1] += m.colptr[c];
  return m;
}

// ---------------------------------------------------------------------------------------------
// Problem + Linearizer
// ---------------------------------------------------------------------------------------------
struct KeyHelper {  // linearization_dense_key_helper_t (lcmtypes/symforce.lcm:67-91)
  int32_t factor_offset, combined_offset, tangent_dim;
  std::vector<int32_t> hessian_storage_col_starts;
  std::vector<int32_t> jacobian_storage_col_starts;
};
struct FactorHelper {
  int32_t residual_dim, combined_residual_offset;
  std::vector<KeyHelper> key_helpers;
};

struct Problem {
  std::vector<KeyT> keys;
  std::vector<int32_t> types, storage_offset;
  int64_t storage_total = 0;
  std::vector<double> values;
  std::vector<int32_t> ftype;
  std::vector<std::vector<int32_t>> fargs;
  std::vector<int32_t> state_keys;       // key indices in state order
  std::vector<int32_t> state_offset_of;  // per key: tangent offset or -1
  int32_t N = 0, M = 0;
  bool include_jacobians = false;
  std::vector<FactorHelper> helpers;
  Csc H, J;
};

static std::vector<int32_t> compute_keys_to_optimize(const Problem& p) {
  std::vector<char> used(p.keys.size(), 0);
  for (size_t f = 0; f < p.ftype.size(); ++f) {
    const FactorSpec s = factor_spec(p.ftype[f]);
    for (int a = 0; a < s.nargs; ++a)
      if (s.lin_mask & (1 << a)) used[p.fargs[f][a]] = 1;
  }
  std::vector<int32_t> out;
  for (size_t k = 0; k < used.size(); ++k)
    if (used[k]) out.push_back((int32_t)k);
  std::sort(out.begin(), out.end(),
            [&](int32_t a, int32_t b) { return lexical_less(p.keys[a], p.keys[b]); });
  return out;
}

static std::vector<const double*> factor_arg_ptrs(const Problem& p, int f, const double* vals) {
  std::vector<const double*> a;
  for (int32_t k : p.fargs[f]) a.push_back(vals + p.storage_offset[k]);
  return a;
}

// BuildInitialLinearization (opt/linearizer.cc:148-356): state index, pattern and storage maps.
static void build_structure(Problem& p) {
  p.state_offset_of.assign(p.keys.size(), -1);
  int32_t off = 0;
  for (int32_t k : p.state_keys) {
    p.state_offset_of[k] = off;
    off += tangent_dim(p.types[k]);
  }
  p.N = off;
  std::vector<char> touched(p.keys.size(), 0);
  std::vector<Triplet> htrip, jtrip;
  for (int32_t i = 0; i < p.N; ++i) htrip.push_back({i, i, 0.0});
  int32_t res_off = 0;
  p.helpers.clear();
  for (size_t f = 0; f < p.ftype.size(); ++f) {
    const FactorSpec s = factor_spec(p.ftype[f]);
    FactorHelper h;
    h.residual_dim = s.res_dim;
    h.combined_residual_offset = res_off;
    res_off += s.res_dim;
    int32_t foff = 0;  // FactorOffsets (linearizer_utils.h:218-252)
    for (int a = 0; a < s.nargs; ++a) {
      if (!(s.lin_mask & (1 << a))) continue;
      const int32_t k = p.fargs[f][a];
      const int32_t td = tangent_dim(p.types[k]);
      if (p.state_offset_of[k] >= 0) {
        touched[k] = 1;
        h.key_helpers.push_back({foff, p.state_offset_of[k], td, {}, {}});
      }
      foff += td;
    }
    if (h.key_helpers.empty())
      std::fprintf(stderr, "oracle: factor %zu touches no optimized keys\n", f);
    // UpdateFromDenseFactorIntoTripletLists (linearizer.cc:475-538), structure only
    for (size_t ki = 0; ki < h.key_helpers.size(); ++ki) {
      const KeyHelper& a = h.key_helpers[ki];
      for (int r = 0; r < h.residual_dim; ++r)
        for (int c = 0; c < a.tangent_dim; ++c)
          jtrip.push_back({h.combined_residual_offset + r, a.combined_offset + c, 0.0});
      for (int r = 0; r < a.tangent_dim; ++r)
        for (int c = 0; c <= r; ++c)
          htrip.push_back({a.combined_offset + r, a.combined_offset + c, 0.0});
      for (size_t kj = 0; kj < ki; ++kj) {
        const KeyHelper& b = h.key_helpers[kj];
        if (a.combined_offset > b.combined_offset) {
          for (int r = 0; r < a.tangent_dim; ++r)
            for (int c = 0; c < b.tangent_dim; ++c)
              htrip.push_back({a.combined_offset + r, b.combined_offset + c, 0.0});
        } else {
          for (int r = 0; r < b.tangent_dim; ++r)
            for (int c = 0; c < a.tangent_dim; ++c)
              htrip.push_back({b.combined_offset + r, a.combined_offset + c, 0.0});
        }
      }
    }
    p.helpers.push_back(std::move(h));
  }
  for (int32_t k : p.state_keys)
    if (!touched[k])
      throw std::runtime_error("Key is in the state vector but is not optimized by any factor.");
  p.M = res_off;
  p.H = csc_from_triplets(p.N, p.N, htrip);
  if (p.include_jacobians) p.J = csc_from_triplets(p.M, p.N, jtrip);

  // CoordsToStorageOffset + ComputeKeyHelperSparseColOffsets (linearizer_utils.h:82-143)
  auto at = [](const Csc& m, int32_t row, int32_t col) -> int32_t {
    const int32_t* b = m.rowind.data() + m.colptr[col];
    const int32_t* e = m.rowind.data() + m.colptr[col + 1];
    const int32_t* it = std::lower_bound(b, e, row);
    if (it == e || *it != row) throw std::out_of_range("Key not in CoordsToStorageMap");
    return (int32_t)(it - m.rowind.data());
  };
  for (FactorHelper& h : p.helpers) {
    for (size_t ki = 0; ki < h.key_helpers.size(); ++ki) {
      KeyHelper& a = h.key_helpers[ki];
      if (p.include_jacobians)
        for (int c = 0; c < a.tangent_dim; ++c)
          a.jacobian_storage_col_starts.push_back(
              at(p.J, h.combined_residual_offset, a.combined_offset + c));
      for (int c = 0; c < a.tangent_dim; ++c)
        a.hessian_storage_col_starts.push_back(
            at(p.H, a.combined_offset + c, a.combined_offset + c));
      for (size_t kj = 0; kj < ki; ++kj) {
        const KeyHelper& b = h.key_helpers[kj];
        if (b.combined_offset < a.combined_offset) {
          for (int c = 0; c < b.tangent_dim; ++c)
            a.hessian_storage_col_starts.push_back(
                at(p.H, a.combined_offset, b.combined_offset + c));
        } else {
          for (int c = 0; c < a.tangent_dim; ++c)
            a.hessian_storage_col_starts.push_back(
                at(p.H, b.combined_offset, a.combined_offset + c));
        }
      }
    }
  }
}

// Relinearize steady state (linearizer.cc:57-112) + UpdateFromLinearizedDenseFactorIntoSparse
// (linearizer.cc:358-433).
struct LinearizationOut {
  std::vector<double> residual, hval, rhs, jval;
};
static void relinearize(const Problem& p, const double* vals, LinearizationOut& out) {
  out.residual.assign(p.M, 0.0);
  out.rhs.assign(p.N, 0.0);
  out.hval.assign(p.H.nnz(), 0.0);
  if (p.include_jacobians) out.jval.assign(p.J.nnz(), 0.0);
  LinearizedFactor lf;
  for (size_t f = 0; f < p.ftype.size(); ++f) {
    linearize_factor(p.ftype[f], factor_arg_ptrs(p, (int)f, vals), lf);
    const FactorHelper& h = p.helpers[f];
    const int T = lf.tdim;
    for (int i = 0; i < h.residual_dim; ++i) out.residual[h.combined_residual_offset + i] = lf.residual[i];
    for (size_t ki = 0; ki < h.key_helpers.size(); ++ki) {
      const KeyHelper& a = h.key_helpers[ki];
      if (p.include_jacobians)
        for (int c = 0; c < a.tangent_dim; ++c)
          for (int r = 0; r < h.residual_dim; ++r)
            out.jval[a.jacobian_storage_col_starts[c] + r] =
                lf.jacobian[(a.factor_offset + c) * h.residual_dim + r];
      for (int i = 0; i < a.tangent_dim; ++i) out.rhs[a.combined_offset + i] += lf.rhs[a.factor_offset + i];
      size_t cs = 0;
      for (int col = 0; col < a.tangent_dim; ++col) {
        const int32_t start = a.hessian_storage_col_starts[cs++];
        for (int r = 0; r < a.tangent_dim - col; ++r)
          out.hval[start + r] += lf.hessian[(a.factor_offset + col) * T + a.factor_offset + col + r];
      }
      for (size_t kj = 0; kj < ki; ++kj) {
        const KeyHelper& b = h.key_helpers[kj];
        if (b.combined_offset < a.combined_offset) {
          for (int cj = 0; cj < b.tangent_dim; ++cj) {
            const int32_t start = a.hessian_storage_col_starts[cs++];
            for (int r = 0; r < a.tangent_dim; ++r)
              out.hval[start + r] += lf.hessian[(b.factor_offset + cj) * T + a.factor_offset + r];
          }
        } else {
          for (int ci = 0; ci < a.tangent_dim; ++ci) {
            const int32_t start = a.hessian_storage_col_starts[cs++];
            for (int r = 0; r < b.tangent_dim; ++r)
              out.hval[start + r] += lf.hessian[(b.factor_offset + r) * T + a.factor_offset + ci];
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// SparseCholeskySolver (sparse_cholesky_solver.tcc)
// ---------------------------------------------------------------------------------------------
struct Cholesky {
  int32_t N = 0;
  std::vector<int32_t> perm, inv_perm;  // permutation_ (old->new), inv_permutation_ (new->old)
  Csc Ap;                               // A_permuted_ (upper)
  std::vector<int32_t> visited, parent, nnz_per_col, L_outer, L_inner, L_k_pattern;
  std::vector<double> L_value, D, D_agg;
  bool initialized = false;

  // selfadjointView<Lower>().twistedBy(P) into Upper storage (Eigen permute_symm_to_symm):
  // destination column max(ip,jp), row min(ip,jp), entries in source traversal order.
  void twist(const Csc& A) {
    const bool has_perm = !perm.empty();
    Ap.rows = Ap.cols = N;
    std::vector<int32_t> count(N, 0);
    for (int32_t j = 0; j < N; ++j) {
      const int32_t jp = has_perm ? perm[j] : j;
      for (int32_t q = A.colptr[j]; q < A.colptr[j + 1]; ++q) {
        const int32_t i = A.rowind[q];
        if (i < j) continue;
        const int32_t ip = has_perm ? perm[i] : i;
        count[std::max(ip, jp)]++;
      }
    }
    Ap.colptr.assign(N + 1, 0);
    for (int32_t j = 0; j < N; ++j) Ap.colptr[j + 1] = Ap.colptr[j] + count[j];
    Ap.rowind.assign(Ap.colptr[N], 0);
    Ap.val.assign(Ap.colptr[N], 0.0);
    for (int32_t j = 0; j < N; ++j) count[j] = Ap.colptr[j];
    for (int32_t j = 0; j < N; ++j) {
      for (int32_t q = A.colptr[j]; q < A.colptr[j + 1]; ++q) {
        const int32_t i = A.rowind[q];
        if (i < j) continue;
        const int32_t jp = has_perm ? perm[j] : j;
        const int32_t ip = has_perm ? perm[i] : i;
        const int32_t k = count[std::max(ip, jp)]++;
        Ap.rowind[k] = std::min(ip, jp);
        Ap.val[k] = A.val.empty() ? 0.0 : A.val[q];
      }
    }
  }

  // ComputePermutationMatrix + ComputeSymbolicSparsity (tcc:12-106)
  void analyze(const Csc& A, const int32_t* ordering) {
    N = A.cols;
    if (ordering) {
      inv_perm.assign(ordering, ordering + N);
      perm.assign(N, 0);
      for (int32_t i = 0; i < N; ++i) perm[inv_perm[i]] = i;
    } else {
      perm.clear();
      inv_perm.clear();
    }
    twist(A);
    visited.assign(N, -1);
    parent.assign(N, -1);
    nnz_per_col.assign(N, 0);
    for (int32_t k = 0; k < N; ++k) {
      visited[k] = k;
      for (int32_t q = Ap.colptr[k]; q < Ap.colptr[k + 1]; ++q) {
        int32_t i = Ap.rowind[q];
        if (i >= k) continue;
        while (visited[i] != k) {
          if (parent[i] == -1) parent[i] = k;
          nnz_per_col[i]++;
          visited[i] = k;
          i = parent[i];
        }
      }
    }
    L_outer.assign(N + 1, 0);
    for (int32_t k = 0; k < N; ++k) L_outer[k + 1] = L_outer[k] + nnz_per_col[k];
    L_inner.assign(L_outer[N], 0);
    L_value.assign(L_outer[N], 0.0);
    D.assign(N, 0.0);
    L_k_pattern.assign(N, 0);
    D_agg.assign(N, 0.0);
    initialized = true;
  }

  // Factorize (tcc:108-219): up-looking LDL^T
  void factorize(const Csc& A) {
    twist(A);
    std::fill(nnz_per_col.begin(), nnz_per_col.end(), 0);
    std::fill(D_agg.begin(), D_agg.end(), 0.0);
    for (int32_t k = 0; k < N; ++k) {
      visited[k] = k;
      int32_t top = N;
      for (int32_t q = Ap.colptr[k]; q < Ap.colptr[k + 1]; ++q) {
        int32_t i = Ap.rowind[q];
        if (i > k) continue;
        D_agg[i] += Ap.val[q];
        int32_t depth = 0;
        while (visited[i] != k) {
          L_k_pattern[depth] = i;
          visited[i] = k;
          i = parent[i];
          depth++;
        }
        while (depth > 0) {
          top--;
          depth--;
          L_k_pattern[top] = L_k_pattern[depth];
        }
      }
      double D_k = D_agg[k];
      D_agg[k] = 0.0;
      for (; top < N; ++top) {
        const int32_t i = L_k_pattern[top];
        const double D_agg_i = D_agg[i];
        const double L_ki = D_agg_i / D[i];
        const int32_t ptr_start = L_outer[i];
        const int32_t ptr_end = ptr_start + nnz_per_col[i];
        D_agg[i] = 0.0;
        int32_t ptr;
        for (ptr = ptr_start; ptr < ptr_end; ++ptr) D_agg[L_inner[ptr]] -= L_value[ptr] * D_agg_i;
        L_inner[ptr] = k;
        L_value[ptr] = L_ki;
        D_k -= L_ki * D_agg_i;
        nnz_per_col[i] += 1;
      }
      D[k] = D_k;
    }
  }

  // SolveInPlace (tcc:230-259) for one column
  void solve(double* x) const {
    std::vector<double> y(N);
    if (!perm.empty()) {
      for (int32_t i = 0; i < N; ++i) y[perm[i]] = x[i];
    } else {
      std::copy(x, x + N, y.begin());
    }
    for (int32_t j = 0; j < N; ++j)  // unit lower forward solve, column oriented
      for (int32_t q = L_outer[j]; q < L_outer[j + 1]; ++q) y[L_inner[q]] -= L_value[q] * y[j];
    for (int32_t j = 0; j < N; ++j) y[j] /= D[j];
    for (int32_t j = N - 1; j >= 0; --j)  // L^T back solve
      for (int32_t q = L_outer[j]; q < L_outer[j + 1]; ++q) y[j] -= L_value[q] * y[L_inner[q]];
    if (!inv_perm.empty()) {
      for (int32_t i = 0; i < N; ++i) x[inv_perm[i]] = y[i];
    } else {
      std::copy(y.begin(), y.end(), x);
    }
  }
};

// ---------------------------------------------------------------------------------------------
// SparseSchurSolver (sparse_schur_solver.tcc)
// ---------------------------------------------------------------------------------------------
struct Schur {
  int32_t total = 0, B = 0, C = 0;
  struct CBlock {
    int32_t start, dim;
  };
  std::vector<CBlock> blocks;
  std::vector<int32_t> block_of;  // per C column: block index
  std::vector<std::vector<double>> Cinv;  // dense dim x dim per block
  // E^T = A[B:, :B] as per-C-row lists of (B col, value position in A)
  std::vector<std::vector<std::pair<int32_t, int32_t>>> et_rows;
  Csc S;
  std::vector<int64_t> s_pos;  // dense (col*B+row) -> position in S, or -1
  Cholesky s_solver;
  std::vector<int32_t> s_ordering;
  bool has_s_ordering = false;
  std::vector<double> a_vals;

  void analyze(const Csc& A, int32_t c_dim, const int32_t* ordering) {
    total = A.cols;
    C = c_dim;
    B = total - C;
    blocks.clear();
    block_of.assign(C, -1);
    bool in_block = false;
    for (int32_t col = B; col < total; ++col) {  // tcc:25-80
      int32_t start_row = -1, prev_row = -1;
      for (int32_t q = A.colptr[col]; q < A.colptr[col + 1]; ++q) {
        const int32_t r = A.rowind[q];
        if (start_row == -1) start_row = r;
        if (prev_row != -1 && r != prev_row + 1)
          throw std::runtime_error("Submatrix C of A is not block diagonal");
        prev_row = r;
      }
      if (!(start_row != -1 && start_row == col)) throw std::runtime_error("C column invalid");
      const int32_t nz = prev_row - start_row + 1;
      if (in_block) {
        const CBlock& b = blocks.back();
        const int32_t bo = col - b.start;
        if (nz != b.dim - bo) throw std::runtime_error("C block shape mismatch");
        if (bo == b.dim - 1) in_block = false;
      } else {
        blocks.push_back({col, nz});
        if (nz > 1) in_block = true;
      }
      block_of[col - B] = (int32_t)blocks.size() - 1;
    }
    // E^T rows: entries of A with row >= B and col < B
    et_rows.assign(C, {});
    for (int32_t col = 0; col < B; ++col)
      for (int32_t q = A.colptr[col]; q < A.colptr[col + 1]; ++q)
        if (A.rowind[q] >= B) et_rows[A.rowind[q] - B].push_back({col, q});
    // S pattern = lower(B) U lower(E^T C^-1 E) structural: marked in a dense B x B mask and
    // read out column by column, which gives the sorted, duplicate-free CSC that setFromTriplets
    // would (B is the small camera-side dimension; the dense position table below is B x B too)
    std::vector<uint8_t> mask((size_t)B * B, 0);
    for (int32_t col = 0; col < B; ++col)
      for (int32_t q = A.colptr[col]; q < A.colptr[col + 1]; ++q)
        if (A.rowind[q] < B) mask[(size_t)col * B + A.rowind[q]] = 1;
    for (const CBlock& b : blocks) {
      std::vector<int32_t> cols;
      for (int32_t r = 0; r < b.dim; ++r)
        for (auto& e : et_rows[b.start - B + r]) cols.push_back(e.first);
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      for (size_t x = 0; x < cols.size(); ++x)
        for (size_t y = 0; y <= x; ++y) mask[(size_t)cols[y] * B + cols[x]] = 1;
    }
    S = Csc{};
    S.rows = S.cols = B;
    S.colptr.assign(B + 1, 0);
    for (int32_t col = 0; col < B; ++col) {
      const uint8_t* m = mask.data() + (size_t)col * B;
      for (int32_t row = col; row < B; ++row)
        if (m[row]) S.rowind.push_back(row);
      S.colptr[col + 1] = (int32_t)S.rowind.size();
    }
    S.val.assign(S.rowind.size(), 0.0);
    s_pos.assign((int64_t)B * B, -1);
    for (int32_t col = 0; col < B; ++col)
      for (int32_t q = S.colptr[col]; q < S.colptr[col + 1]; ++q)
        s_pos[(int64_t)col * B + S.rowind[q]] = q;
    has_s_ordering = ordering != nullptr;
    if (ordering) s_ordering.assign(ordering, ordering + B);
    Cinv.assign(blocks.size(), {});
  }

  // Factorize (tcc:101-139)
  void factorize(const Csc& A) {
    a_vals = A.val;
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
      const CBlock& b = blocks[bi];
      const int d = b.dim;
      std::vector<double> M(d * d, 0.0), Lc(d * d, 0.0);
      for (int c = 0; c < d; ++c)
        for (int32_t q = A.colptr[b.start + c]; q < A.colptr[b.start + c + 1]; ++q) {
          const int r = A.rowind[q] - b.start;
          M[c * d + r] = A.val[q];
        }
      // dense LLT (lower) then solve for identity
      for (int j = 0; j < d; ++j) {
        double s = M[j * d + j];
        for (int k = 0; k < j; ++k) s -= Lc[k * d + j] * Lc[k * d + j];
        const double ljj = std::sqrt(s);
        Lc[j * d + j] = ljj;
        for (int i = j + 1; i < d; ++i) {
          double v = M[j * d + i];
          for (int k = 0; k < j; ++k) v -= Lc[k * d + i] * Lc[k * d + j];
          Lc[j * d + i] = v / ljj;
        }
      }
      std::vector<double> inv(d * d, 0.0);
      for (int c = 0; c < d; ++c) {
        std::vector<double> y(d, 0.0);
        for (int i = 0; i < d; ++i) {
          double v = (i == c) ? 1.0 : 0.0;
          for (int k = 0; k < i; ++k) v -= Lc[k * d + i] * y[k];
          y[i] = v / Lc[i * d + i];
        }
        for (int i = d - 1; i >= 0; --i) {
          double v = y[i];
          for (int k = i + 1; k < d; ++k) v -= Lc[i * d + k] * y[k];
          y[i] = v / Lc[i * d + i];
        }
        for (int i = 0; i < d; ++i) inv[c * d + i] = y[i];
      }
      Cinv[bi] = inv;
    }
    std::fill(S.val.begin(), S.val.end(), 0.0);
    for (int32_t col = 0; col < B; ++col)
      for (int32_t q = A.colptr[col]; q < A.colptr[col + 1]; ++q)
        if (A.rowind[q] < B) S.val[s_pos[(int64_t)col * B + A.rowind[q]]] += A.val[q];
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
      const CBlock& b = blocks[bi];
      const int d = b.dim;
      // Y = C^-1 E for this block: rows r (block-local), per B column
      std::vector<int32_t> cols;
      for (int r = 0; r < d; ++r)
        for (auto& e : et_rows[b.start - B + r]) cols.push_back(e.first);
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      const size_t nc = cols.size();
      std::vector<double> E(d * nc, 0.0);  // E^T block: d rows x nc cols
      for (int r = 0; r < d; ++r)
        for (auto& e : et_rows[b.start - B + r]) {
          const size_t ci = std::lower_bound(cols.begin(), cols.end(), e.first) - cols.begin();
          E[ci * d + r] = A.val[e.second];
        }
      std::vector<double> Y(d * nc, 0.0);
      for (size_t c = 0; c < nc; ++c)
        for (int r = 0; r < d; ++r) {
          double v = 0;
          for (int k = 0; k < d; ++k) v += Cinv[bi][k * d + r] * E[c * d + k];
          Y[c * d + r] = v;
        }
      for (size_t x = 0; x < nc; ++x)
        for (size_t y = 0; y <= x; ++y) {
          double v = 0;
          for (int k = 0; k < d; ++k) v += E[x * d + k] * Y[y * d + k];
          S.val[s_pos[(int64_t)cols[y] * B + cols[x]]] -= v;
        }
    }
    if (!s_solver.initialized) s_solver.analyze(S, has_s_ordering ? s_ordering.data() : nullptr);
    s_solver.factorize(S);
  }

  // Solve (tcc:143-163)
  void solve(const double* rhs, double* out) const {
    std::vector<double> v(rhs, rhs + B), w(rhs + B, rhs + total);
    std::vector<double> ciw(C, 0.0);
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
      const CBlock& b = blocks[bi];
      const int d = b.dim;
      for (int r = 0; r < d; ++r) {
        double acc = 0;
        for (int k = 0; k < d; ++k) acc += Cinv[bi][k * d + r] * w[b.start - B + k];
        ciw[b.start - B + r] = acc;
      }
    }
    for (int32_t r = 0; r < C; ++r)
      for (auto& e : et_rows[r]) v[e.first] -= a_vals[e.second] * ciw[r];
    s_solver.solve(v.data());  // y
    std::vector<double> t = w;
    for (int32_t r = 0; r < C; ++r)
      for (auto& e : et_rows[r]) t[r] -= a_vals[e.second] * v[e.first];
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
      const CBlock& b = blocks[bi];
      const int d = b.dim;
      for (int r = 0; r < d; ++r) {
        double acc = 0;
        for (int k = 0; k < d; ++k) acc += Cinv[bi][k * d + r] * t[b.start - B + k];
        out[b.start + r] = acc;
      }
    }
    std::copy(v.begin(), v.end(), out);
  }
};

// ---------------------------------------------------------------------------------------------
// Levenberg-Marquardt (levenberg_marquardt_solver.tcc + levenberg_marquardt_state.h)
// ---------------------------------------------------------------------------------------------
struct StateBlock {
  std::vector<double> values;
  LinearizationOut lin;
  bool linearized = false;
  bool have_cached_error = false;
  double cached_error = 0;
  double Error() {
    if (!have_cached_error) {
      double s = 0;
      for (double r : lin.residual) s += r * r;
      cached_error = 0.5 * s;
      have_cached_error = true;
    }
    return cached_error;
  }
};

struct LinearSolver {
  int kind = ORACLE_SOLVER_CHOLESKY;
  std::vector<int32_t> ordering;
  bool has_ordering = false;
  int32_t c_dim = 0;
  Cholesky chol;
  Schur schur;
  void analyze(const Csc& H) {
    if (kind == ORACLE_SOLVER_CHOLESKY)
      chol.analyze(H, has_ordering ? ordering.data() : nullptr);
    else
      schur.analyze(H, c_dim, has_ordering ? ordering.data() : nullptr);
  }
  void factorize(const Csc& H) {
    if (kind == ORACLE_SOLVER_CHOLESKY)
      chol.factorize(H);
    else
      schur.factorize(H);
  }
  void solve(const std::vector<double>& b, std::vector<double>& x) {
    x = b;
    if (kind == ORACLE_SOLVER_CHOLESKY)
      chol.solve(x.data());
    else
      schur.solve(b.data(), x.data());
  }
};

struct LM {
  const Problem* prob = nullptr;
  symhip_optimizer_params p{};
  double epsilon = 0;
  StateBlock blocks[3];
  int init_idx = 0, new_idx = 1, best_idx = 0, free_idx = 2;
  bool initialized_[3] = {false, false, false};
  bool best_valid = false;
  double lambda = 0, nu = 0;
  int iteration = -1;
  bool have_max_diagonal = false, have_last_update = false, solver_analyzed = false;
  std::vector<double> max_diagonal, damping, undamped, update, last_update;
  LinearSolver solver;
  std::vector<std::vector<double>>* updates_log = nullptr;

  StateBlock& Init() { return blocks[init_idx]; }
  StateBlock& New() { return blocks[new_idx]; }
  StateBlock& Best() { return blocks[best_idx]; }
  void SwapNewAndInit() { std::swap(init_idx, new_idx); }
  void SetBestToNew() {
    best_valid = true;
    if (best_idx == new_idx) return;
    if (best_idx != init_idx) free_idx = best_idx;
    best_idx = new_idx;
  }
  void SetBestToInit() {
    best_valid = true;
    if (best_idx == init_idx) return;
    if (best_idx != new_idx) free_idx = best_idx;
    best_idx = init_idx;
  }
  void SetInitToNotBest() {
    if (best_idx != init_idx) return;
    init_idx = free_idx;
    free_idx = best_idx;
  }

  void Relinearize(StateBlock& b) {
    relinearize(*prob, b.values.data(), b.lin);
    b.linearized = true;
    b.have_cached_error = false;
  }

  void Reset(const std::vector<double>& values) {  // levenberg_marquardt_solver.h:163-168
    lambda = p.initial_lambda;
    nu = p.dynamic_lambda_update_beta;
    iteration = -1;
    ResetState(values);
  }
  void RelaxDampingToInitial() {  // levenberg_marquardt_solver.h:171-174
    lambda = std::min(lambda, p.initial_lambda);
    nu = p.dynamic_lambda_update_beta;
  }
  void ResetState(const std::vector<double>& values) {  // levenberg_marquardt_solver.h:178-183
    have_max_diagonal = false;
    have_last_update = false;
    // LevenbergMarquardtState::ResetValues + StateBase::Reset
    New().values = values;
    initialized_[new_idx] = true;
    initialized_[init_idx] = false;
    initialized_[free_idx] = false;
    for (auto& b : blocks) {
      b.linearized = false;
      b.have_cached_error = false;
    }
    best_valid = false;
  }

  // UpdateNewFromInitImpl (state.h:217-239) + Values::Retract (values.cc:314-327)
  void UpdateNewFromInit() {
    auto& nv = New().values;
    const auto& iv = Init().values;
    if (nv.empty() || !initialized_[new_idx]) {
      nv = iv;
      initialized_[new_idx] = true;
    } else {
      for (int32_t k : prob->state_keys) {
        const int32_t o = prob->storage_offset[k];
        const int sd = storage_dim(prob->types[k]);
        std::copy(iv.begin() + o, iv.begin() + o + sd, nv.begin() + o);
      }
    }
    int32_t t = 0;
    for (int32_t k : prob->state_keys) {
      double* x = nv.data() + prob->storage_offset[k];
      const int type = prob->types[k];
      const double* d = update.data() + t;
      if (type == SYMHIP_TYPE_POSE3) {
        double out[7];
        pose3_retract(x, d, epsilon, out);
        std::copy(out, out + 7, x);
      } else if (type == SYMHIP_TYPE_ROT3) {
        double out[4];
        rot3_retract(x, d, epsilon, out);
        std::copy(out, out + 4, x);
      } else {
        for (int i = 0; i < tangent_dim(type); ++i) x[i] += d[i];
      }
      t += tangent_dim(type);
    }
  }

  static void get_diag(const Problem& pr, const std::vector<double>& hval, std::vector<double>& d) {
    d.resize(pr.N);
    for (int32_t j = 0; j < pr.N; ++j) d[j] = hval[pr.H.colptr[j]];  // diagonal first (rows sorted)
  }
  static void set_diag(const Problem& pr, std::vector<double>& hval, const std::vector<double>& d) {
    for (int32_t j = 0; j < pr.N; ++j) hval[pr.H.colptr[j]] = d[j];
  }

  // DampHessian (tcc:22-54)
  void DampHessian(std::vector<double>& hval) {
    const int32_t N = prob->N;
    get_diag(*prob, hval, undamped);
    damping.assign(N, 0.0);
    if (p.use_diagonal_damping) {
      if (p.keep_max_diagonal_damping) {
        if (!have_max_diagonal) {
          max_diagonal.resize(N);
          for (int32_t i = 0; i < N; ++i) max_diagonal[i] = std::max(undamped[i], p.diagonal_damping_min);
        } else {
          for (int32_t i = 0; i < N; ++i) max_diagonal[i] = std::max(max_diagonal[i], undamped[i]);
        }
        have_max_diagonal = true;
        for (int32_t i = 0; i < N; ++i) damping[i] = max_diagonal[i] * lambda;
      } else {
        for (int32_t i = 0; i < N; ++i) damping[i] = std::max(undamped[i], p.diagonal_damping_min) * lambda;
      }
    }
    if (p.use_unit_damping)
      for (int32_t i = 0; i < N; ++i) damping[i] += lambda;
    for (int32_t j = 0; j < N; ++j) hval[prob->H.colptr[j]] += damping[j];
  }

  // Iterate (tcc:138-343)
  std::optional<std::pair<int, int>> Iterate(std::vector<symhip_iteration_stats>& stats,
                                             int32_t& best_index) {
    SwapNewAndInit();  // state_.Step()
    iteration++;
    if (!Init().linearized) {
      Relinearize(Init());
      SetBestToInit();
    }
    if (iteration == 0) {
      symhip_iteration_stats s{};
      s.iteration = -1;
      s.new_error = Init().Error();
      s.current_lambda = lambda;
      stats.push_back(s);
      if (!std::isfinite(Init().Error()))
        return std::make_pair((int)SYMHIP_STATUS_FAILED, (int)SYMHIP_FAILURE_INITIAL_ERROR_NOT_FINITE);
    }
    Csc Hm = prob->H;
    if (!solver_analyzed) {
      solver.analyze(Hm);
      solver_analyzed = true;
    }
    DampHessian(Init().lin.hval);
    Hm.val = Init().lin.hval;
    solver.factorize(Hm);
    std::vector<double> x;
    solver.solve(Init().lin.rhs, x);
    update.resize(x.size());
    for (size_t i = 0; i < x.size(); ++i) update[i] = -x[i];
    set_diag(*prob, Init().lin.hval, undamped);
    if (updates_log) updates_log->push_back(update);
    UpdateNewFromInit();
    Relinearize(New());
    const double e0 = Init().Error();
    const double new_error = New().Error();
    const double relative_reduction = (e0 - new_error) / (e0 + epsilon);
    // LinearDeltaError (linearization.h:62-67)
    double lde = 0;
    for (size_t i = 0; i < update.size(); ++i)
      lde += update[i] * (Init().lin.rhs[i] - damping[i] * update[i]);
    const double new_error_linear = e0 + 0.5 * lde;
    const double gain_ratio = (e0 - new_error) / (e0 - new_error_linear);
    symhip_iteration_stats s{};
    s.iteration = iteration;
    s.current_lambda = lambda;
    s.new_error = new_error;
    s.new_error_linear = new_error_linear;
    s.relative_reduction = relative_reduction;
    stats.push_back(s);
    const size_t rec = stats.size() - 1;
    std::optional<std::pair<int, int>> status;
    if (relative_reduction > -p.early_exit_min_reduction / 10 &&
        relative_reduction < p.early_exit_min_reduction) {
      status = std::make_pair((int)SYMHIP_STATUS_SUCCESS, 0);
    } else if (new_error < p.early_exit_min_absolute_error) {
      status = std::make_pair((int)SYMHIP_STATUS_SUCCESS, 0);
    }
    bool accept = relative_reduction > 0;
    double angle = 0;
    if (p.enable_bold_updates && have_last_update && !accept) {
      // last_update_.normalized().dot(update_.stableNormalized())
      double nl = 0, nu_ = 0, dot = 0;
      for (size_t i = 0; i < update.size(); ++i) {
        nl += last_update[i] * last_update[i];
        nu_ += update[i] * update[i];
        dot += last_update[i] * update[i];
      }
      angle = (nl > 0 && nu_ > 0) ? dot / (std::sqrt(nl) * std::sqrt(nu_)) : 0.0;
      accept = ((1 - angle) * (1 - angle) * new_error) <= Best().Error();
    }
    if (!accept && lambda >= p.lambda_upper_bound)
      status = std::make_pair((int)SYMHIP_STATUS_FAILED, (int)SYMHIP_FAILURE_LAMBDA_OUT_OF_BOUNDS);
    if (!accept) {
      if (p.lambda_update_type == SYMHIP_LAMBDA_STATIC) {
        lambda *= p.lambda_up_factor;
      } else if (p.lambda_update_type == SYMHIP_LAMBDA_DYNAMIC) {
        lambda *= nu;
        nu *= 2;
      } else {
        throw std::runtime_error("Invalid lambda update type");
      }
      SwapNewAndInit();
    } else {
      if (p.lambda_update_type == SYMHIP_LAMBDA_STATIC) {
        lambda *= p.lambda_down_factor;
      } else if (p.lambda_update_type == SYMHIP_LAMBDA_DYNAMIC) {
        lambda *= std::max(1.0 / p.dynamic_lambda_update_gamma,
                           1.0 - (p.dynamic_lambda_update_beta - 1) *
                                     std::pow(2 * gain_ratio - 1, (double)p.dynamic_lambda_update_p));
        nu = 2;
      } else {
        throw std::runtime_error("Invalid lambda update type");
      }
      have_last_update = true;
      last_update = update;
      if (New().Error() <= Best().Error()) {
        SetBestToNew();
        best_index = (int32_t)rec;
      }
      SetInitToNotBest();
    }
    lambda = std::min(std::max(lambda, p.lambda_lower_bound), p.lambda_upper_bound);
    stats[rec].update_angle_change = angle;
    stats[rec].update_accepted = accept ? 1 : 0;
    return status;
  }
};

}  // namespace oracle

using namespace oracle;

struct oracle_problem {
  Problem p;
  bool built = false;
};
struct oracle_cholesky {
  Cholesky c;
  Csc A;
};
struct oracle_schur {
  Schur s;
  Csc A;
};

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

#define ORACLE_TRY try {
#define ORACLE_CATCH(ret)         \
  }                               \
  catch (const std::exception& e) { \
    g_error = e.what();           \
    return ret;                   \
  }

oracle_problem* oracle_problem_create(const symhip_problem_desc* d, int include_jacobians) {
  ORACLE_TRY
  auto* op = new oracle_problem();
  Problem& p = op->p;
  p.include_jacobians = include_jacobians != 0;
  int64_t off = 0;
  for (int32_t k = 0; k < d->n_keys; ++k) {
    p.keys.push_back({d->key_letters[k], d->key_subs[k], d->key_supers[k]});
    p.types.push_back(d->key_types[k]);
    p.storage_offset.push_back((int32_t)off);
    off += storage_dim(d->key_types[k]);
  }
  p.storage_total = off;
  p.values.assign(d->values, d->values + off);
  for (int32_t f = 0; f < d->n_factors; ++f) {
    p.ftype.push_back(d->factor_types[f]);
    std::vector<int32_t> a(d->factor_args + d->factor_arg_ptr[f],
                           d->factor_args + d->factor_arg_ptr[f + 1]);
    const FactorSpec s = factor_spec(d->factor_types[f]);
    if ((int)a.size() != s.nargs) throw std::runtime_error("factor arity mismatch");
    for (int i = 0; i < s.nargs; ++i)
      if (p.types[a[i]] != s.arg_types[i]) throw std::runtime_error("factor argument type mismatch");
    p.fargs.push_back(a);
  }
  if (d->n_opt_keys > 0)
    p.state_keys.assign(d->opt_keys, d->opt_keys + d->n_opt_keys);
  else
    p.state_keys = compute_keys_to_optimize(p);
  build_structure(p);
  op->built = true;
  return op;
  ORACLE_CATCH(nullptr)
}

void oracle_problem_destroy(oracle_problem* p) { delete p; }

int oracle_dims(oracle_problem* op, int64_t* out) {
  out[0] = op->p.N;
  out[1] = op->p.M;
  out[2] = op->p.H.nnz();
  out[3] = op->p.J.nnz();
  return 0;
}
int oracle_hessian_structure(oracle_problem* op, int32_t* colptr, int32_t* rowind) {
  std::copy(op->p.H.colptr.begin(), op->p.H.colptr.end(), colptr);
  std::copy(op->p.H.rowind.begin(), op->p.H.rowind.end(), rowind);
  return 0;
}
int oracle_jacobian_structure(oracle_problem* op, int32_t* colptr, int32_t* rowind) {
  std::copy(op->p.J.colptr.begin(), op->p.J.colptr.end(), colptr);
  std::copy(op->p.J.rowind.begin(), op->p.J.rowind.end(), rowind);
  return 0;
}
int oracle_storage_offsets(oracle_problem* op, int32_t* offsets) {
  std::copy(op->p.storage_offset.begin(), op->p.storage_offset.end(), offsets);
  return 0;
}
int oracle_state_keys(oracle_problem* op, int32_t* key_index, int32_t* tangent_offset) {
  for (size_t i = 0; i < op->p.state_keys.size(); ++i) {
    key_index[i] = op->p.state_keys[i];
    tangent_offset[i] = op->p.state_offset_of[op->p.state_keys[i]];
  }
  return (int)op->p.state_keys.size();
}
int64_t oracle_hessian_col_starts(oracle_problem* op, int32_t* out) {
  int64_t n = 0;
  for (const auto& h : op->p.helpers)
    for (const auto& k : h.key_helpers)
      for (int32_t v : k.hessian_storage_col_starts) {
        if (out) out[n] = v;
        n++;
      }
  return n;
}

int oracle_linearize(oracle_problem* op, const double* values, double* residual,
                     double* hval, double* rhs, double* jval) {
  ORACLE_TRY
  LinearizationOut out;
  relinearize(op->p, values, out);
  if (residual) std::copy(out.residual.begin(), out.residual.end(), residual);
  if (hval) std::copy(out.hval.begin(), out.hval.end(), hval);
  if (rhs) std::copy(out.rhs.begin(), out.rhs.end(), rhs);
  if (jval && op->p.include_jacobians) std::copy(out.jval.begin(), out.jval.end(), jval);
  return 0;
  ORACLE_CATCH(-1)
}

static Csc make_csc(int32_t n, const int32_t* colptr, const int32_t* rowind) {
  Csc A;
  A.rows = A.cols = n;
  A.colptr.assign(colptr, colptr + n + 1);
  A.rowind.assign(rowind, rowind + colptr[n]);
  A.val.assign(colptr[n], 0.0);
  return A;
}

oracle_cholesky* oracle_cholesky_create(int32_t n, const int32_t* colptr, const int32_t* rowind,
                                        const int32_t* ordering) {
  ORACLE_TRY
  auto* c = new oracle_cholesky();
  c->A = make_csc(n, colptr, rowind);
  c->c.analyze(c->A, ordering);
  return c;
  ORACLE_CATCH(nullptr)
}
void oracle_cholesky_destroy(oracle_cholesky* c) { delete c; }
int oracle_cholesky_factorize(oracle_cholesky* c, const double* values) {
  std::copy(values, values + c->A.nnz(), c->A.val.begin());
  c->c.factorize(c->A);
  return 0;
}
int oracle_cholesky_solve(oracle_cholesky* c, const double* b, double* x, int32_t nrhs) {
  const int32_t n = c->c.N;
  for (int32_t k = 0; k < nrhs; ++k) {
    std::copy(b + (int64_t)k * n, b + (int64_t)(k + 1) * n, x + (int64_t)k * n);
    c->c.solve(x + (int64_t)k * n);
  }
  return 0;
}
int64_t oracle_cholesky_nnz(oracle_cholesky* c) { return c->c.L_outer.back(); }
int oracle_cholesky_factor(oracle_cholesky* c, int32_t* colptr, int32_t* rowind, double* lvals,
                           double* d, int32_t* parent) {
  const Cholesky& ch = c->c;
  if (colptr) std::copy(ch.L_outer.begin(), ch.L_outer.end(), colptr);
  if (rowind) std::copy(ch.L_inner.begin(), ch.L_inner.end(), rowind);
  if (lvals) std::copy(ch.L_value.begin(), ch.L_value.end(), lvals);
  if (d) std::copy(ch.D.begin(), ch.D.end(), d);
  if (parent) std::copy(ch.parent.begin(), ch.parent.end(), parent);
  return 0;
}

oracle_schur* oracle_schur_create(int32_t n, const int32_t* colptr, const int32_t* rowind,
                                  int32_t c_dim, const int32_t* s_ordering) {
  ORACLE_TRY
  auto* s = new oracle_schur();
  s->A = make_csc(n, colptr, rowind);
  s->s.analyze(s->A, c_dim, s_ordering);
  return s;
  ORACLE_CATCH(nullptr)
}
void oracle_schur_destroy(oracle_schur* s) { delete s; }
int oracle_schur_factorize(oracle_schur* s, const double* values) {
  ORACLE_TRY
  std::copy(values, values + s->A.nnz(), s->A.val.begin());
  s->s.factorize(s->A);
  return 0;
  ORACLE_CATCH(-1)
}
int oracle_schur_solve(oracle_schur* s, const double* b, double* x) {
  s->s.solve(b, x);
  return 0;
}
int oracle_schur_dense_s(oracle_schur* s, double* out) {
  const int32_t B = s->s.B;
  std::fill(out, out + (int64_t)B * B, 0.0);
  for (int32_t c = 0; c < B; ++c)
    for (int32_t q = s->s.S.colptr[c]; q < s->s.S.colptr[c + 1]; ++q)
      out[(int64_t)c * B + s->s.S.rowind[q]] = s->s.S.val[q];
  return 0;
}

void oracle_default_params(symhip_optimizer_params* p) {
  *p = symhip_optimizer_params{};
  p->initial_lambda = 1.0;
  p->lambda_lower_bound = 0.0;
  p->lambda_upper_bound = 1000000.0;
  p->lambda_update_type = SYMHIP_LAMBDA_STATIC;
  p->lambda_up_factor = 4.0;
  p->lambda_down_factor = 1 / 4.0;
  p->dynamic_lambda_update_beta = 2.0;
  p->dynamic_lambda_update_gamma = 3.0;
  p->dynamic_lambda_update_p = 3;
  p->use_diagonal_damping = 0;
  p->use_unit_damping = 1;
  p->keep_max_diagonal_damping = 0;
  p->diagonal_damping_min = 1e-6;
  p->iterations = 50;
  p->early_exit_min_reduction = 1e-6;
  p->early_exit_min_absolute_error = 0.0;
  p->enable_bold_updates = 0;
}

static double* g_iter_seconds = nullptr;

int oracle_optimize_timed(oracle_problem* op, const symhip_optimizer_params* params, int solver,
                          const int32_t* ordering, int32_t c_dim, double epsilon,
                          int32_t num_iterations, double* values_inout,
                          symhip_iteration_stats* iters, int32_t max_iters,
                          symhip_optimization_stats* out_stats, double* iter_seconds) {
  g_iter_seconds = iter_seconds;
  const int rc = oracle_optimize(op, params, solver, ordering, c_dim, epsilon, num_iterations,
                                 values_inout, iters, max_iters, out_stats, nullptr);
  g_iter_seconds = nullptr;
  return rc;
}

int oracle_optimize(oracle_problem* op, const symhip_optimizer_params* params, int solver,
                    const int32_t* ordering, int32_t c_dim, double epsilon, int32_t num_iterations,
                    double* values_inout, symhip_iteration_stats* iters, int32_t max_iters,
                    symhip_optimization_stats* out_stats, double* updates_out) {
  ORACLE_TRY
  const Problem& pr = op->p;
  LM lm;
  lm.prob = &pr;
  lm.p = *params;
  lm.epsilon = epsilon;
  lm.solver.kind = solver;
  lm.solver.c_dim = c_dim;
  if (ordering) {
    const int32_t n = solver == ORACLE_SOLVER_SCHUR ? pr.N - c_dim : pr.N;
    lm.solver.ordering.assign(ordering, ordering + n);
    lm.solver.has_ordering = true;
  }
  std::vector<std::vector<double>> ulog;
  if (updates_out) lm.updates_log = &ulog;
  std::vector<double> values(values_inout, values_inout + pr.storage_total);
  int32_t n = num_iterations < 0 ? params->iterations : num_iterations;
  if (n <= 0) throw std::runtime_error("num_iterations must be positive");
  lm.Reset(values);
  std::vector<symhip_iteration_stats> stats;
  int32_t best_index = 0;
  int status = SYMHIP_STATUS_INVALID, reason = 0;
  int i;
  for (i = 0; i < n; ++i) {
    const auto t0 = std::chrono::steady_clock::now();
    auto st = lm.Iterate(stats, best_index);
    if (g_iter_seconds)
      g_iter_seconds[i] =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) {
      status = st->first;
      reason = st->second;
      break;
    }
  }
  if (i == n) {
    status = SYMHIP_STATUS_HIT_ITERATION_LIMIT;
    reason = 0;
  }
  const auto& best = lm.Best().values;  // GetBestValues
  std::copy(best.begin(), best.end(), values_inout);
  out_stats->num_iterations = (int32_t)stats.size();
  out_stats->best_index = best_index;
  out_stats->status = status;
  out_stats->failure_reason = reason;
  for (size_t k = 0; k < stats.size() && (int32_t)k < max_iters; ++k) iters[k] = stats[k];
  if (updates_out) {
    int64_t o = 0;
    for (auto& u : ulog)
      for (double v : u) updates_out[o++] = v;
  }
  return 0;
  ORACLE_CATCH(-1)
}

// GncOptimizer::Optimize (opt/gnc_optimizer.h:49-118) over the Optimizer above; each stage is
// IterateToConvergenceImpl (opt/internal/optimizer_utils.h:36-78).
int oracle_optimize_gnc(oracle_problem* op, const symhip_optimizer_params* params, int solver,
                        double epsilon, int32_t num_iterations, const symhip_gnc_params* g,
                        int32_t mu_key, double* values_inout, symhip_iteration_stats* iters,
                        int32_t max_iters, symhip_optimization_stats* out_stats) {
  ORACLE_TRY
  const Problem& pr = op->p;
  if (mu_key < 0 || mu_key >= (int32_t)pr.storage_offset.size())
    throw std::runtime_error("gnc: mu key index out of range");
  if (pr.types[mu_key] != SYMHIP_TYPE_SCALAR) throw std::runtime_error("gnc: mu must be a Scalar");
  const int64_t mu_off = pr.storage_offset[mu_key];
  LM lm;
  lm.prob = &pr;
  lm.p = *params;
  lm.epsilon = epsilon;
  lm.solver.kind = solver;
  int32_t n = num_iterations < 0 ? params->iterations : num_iterations;
  if (n <= 0) throw std::runtime_error("num_iterations must be positive");
  bool updating = g->mu_initial < g->mu_max && g->mu_step > 0.0;
  std::vector<double> values(values_inout, values_inout + pr.storage_total);
  values[mu_off] = g->mu_initial;
  const double early_exit_min_reduction = params->early_exit_min_reduction;
  if (updating) lm.p.early_exit_min_reduction = g->gnc_update_min_reduction;
  lm.Reset(values);
  std::vector<symhip_iteration_stats> stats;
  int32_t best_index = 0;
  int status = SYMHIP_STATUS_INVALID, reason = 0;
  auto iterate_to_convergence = [&](int32_t count) {
    int i;
    for (i = 0; i < count; ++i) {
      auto st = lm.Iterate(stats, best_index);
      if (st) {
        status = st->first;
        reason = st->second;
        break;
      }
    }
    if (i == count) {
      status = SYMHIP_STATUS_HIT_ITERATION_LIMIT;
      reason = 0;
    }
    values = lm.Best().values;  // GetBestValues
  };
  iterate_to_convergence(n);
  while ((int32_t)stats.size() < n) {
    if (status != SYMHIP_STATUS_SUCCESS) break;
    if (!updating) break;
    values[mu_off] = values[mu_off] + g->mu_step;
    lm.RelaxDampingToInitial();
    if (values[mu_off] >= g->mu_max) {
      values[mu_off] = g->mu_max;
      lm.p.early_exit_min_reduction = early_exit_min_reduction;
      updating = false;
    }
    lm.ResetState(values);  // OptimizeContinue (:121-131)
    iterate_to_convergence(n - (int32_t)stats.size());
  }
  std::copy(values.begin(), values.end(), values_inout);
  out_stats->num_iterations = (int32_t)stats.size();
  out_stats->best_index = best_index;
  out_stats->status = status;
  out_stats->failure_reason = reason;
  for (size_t k = 0; k < stats.size() && (int32_t)k < max_iters; ++k) iters[k] = stats[k];
  return 0;
  ORACLE_CATCH(-1)
}

int oracle_linearize_factor(int32_t type, const double* const* args, double* residual,
                            double* jacobian, double* hessian, double* rhs) {
  ORACLE_TRY
  const FactorSpec s = factor_spec(type);
  std::vector<const double*> a(args, args + s.nargs);
  LinearizedFactor f;
  linearize_factor(type, a, f);
  if (residual) std::copy(f.residual.begin(), f.residual.end(), residual);
  if (jacobian) std::copy(f.jacobian.begin(), f.jacobian.end(), jacobian);
  if (hessian) std::copy(f.hessian.begin(), f.hessian.end(), hessian);
  if (rhs) std::copy(f.rhs.begin(), f.rhs.end(), rhs);
  return 0;
  ORACLE_CATCH(-1)
}

}  // extern "C"
