// Host-side problem model. The pattern and storage maps reproduce the reference exactly:
//   ComputeKeysToOptimize           opt/factor.h:423-449 (lexical Key order)
//   FactorOffsets / key helpers     opt/internal/linearizer_utils.h:218-252
//   hessian_lower pattern           opt/linearizer.cc:173-178, :475-538 (triplets) + setFromTriplets
//   hessian_storage_col_starts      opt/internal/linearizer_utils.h:98-143
// The pattern is produced directly from key-block adjacency instead of sorting triplets: every
// optimized key's lower diagonal block is dense (a factor touching it writes its whole lower
// triangle, opt/linearizer.cc:509-514) and off-diagonal key blocks are dense, so column c of key
// K holds rows c..end(K) followed by all rows of each coupled key after K in state order.
#include "./problem.h"

#include <algorithm>
#include <numeric>
#include <tuple>

namespace symhip {

void build_host_problem(const symhip_problem_desc* d, HostProblem& p) {
  SYMHIP_CHECK(d && d->n_keys >= 0 && d->n_factors >= 0, "invalid problem description");
  p = HostProblem();
  p.n_keys = d->n_keys;
  p.letters.assign(d->key_letters, d->key_letters + d->n_keys);
  p.subs.assign(d->key_subs, d->key_subs + d->n_keys);
  p.supers.assign(d->key_supers, d->key_supers + d->n_keys);
  p.types.assign(d->key_types, d->key_types + d->n_keys);
  int64_t off = 0;
  p.storage_offset.resize(p.n_keys);
  for (int32_t k = 0; k < p.n_keys; ++k) {
    p.storage_offset[k] = (int32_t)off;
    off += storage_dim(p.types[k]);
  }
  SYMHIP_CHECK(off < INT32_MAX, "Values storage exceeds 2^31 entries");
  p.storage_total = off;
  p.values.assign(d->values, d->values + off);

  p.n_factors = d->n_factors;
  p.ftype.assign(d->factor_types, d->factor_types + d->n_factors);
  p.fargs_ptr.assign(d->factor_arg_ptr, d->factor_arg_ptr + d->n_factors + 1);
  p.fargs.assign(d->factor_args, d->factor_args + p.fargs_ptr.back());
  for (int32_t f = 0; f < p.n_factors; ++f) {
    const FactorSpec s = factor_spec(p.ftype[f]);
    SYMHIP_CHECK(p.fargs_ptr[f + 1] - p.fargs_ptr[f] == s.nargs,
                 "factor " + std::to_string(f) + " has the wrong number of arguments");
    for (int a = 0; a < s.nargs; ++a) {
      const int32_t k = p.fargs[p.fargs_ptr[f] + a];
      SYMHIP_CHECK(k >= 0 && k < p.n_keys, "factor argument key index out of range");
      SYMHIP_CHECK(p.types[k] == s.arg_types[a],
                   "factor " + std::to_string(f) + " argument " + std::to_string(a) +
                       " has the wrong value type");
    }
  }

  // State keys
  if (d->n_opt_keys > 0) {
    p.state_keys.assign(d->opt_keys, d->opt_keys + d->n_opt_keys);
  } else {
    std::vector<char> used(p.n_keys, 0);
    for (int32_t f = 0; f < p.n_factors; ++f) {
      const FactorSpec s = factor_spec(p.ftype[f]);
      for (int a = 0; a < s.nargs; ++a)
        if (s.lin_mask & (1 << a)) used[p.fargs[p.fargs_ptr[f] + a]] = 1;
    }
    for (int32_t k = 0; k < p.n_keys; ++k)
      if (used[k]) p.state_keys.push_back(k);
    std::sort(p.state_keys.begin(), p.state_keys.end(), [&](int32_t a, int32_t b) {
      return std::make_tuple(p.letters[a], p.subs[a], p.supers[a]) <
             std::make_tuple(p.letters[b], p.subs[b], p.supers[b]);
    });
  }
  const int32_t S = (int32_t)p.state_keys.size();
  p.tan_off.assign(p.n_keys, -1);
  std::vector<int32_t> pos_of(p.n_keys, -1), td_of_pos(S), off_of_pos(S);
  int32_t N = 0;
  for (int32_t i = 0; i < S; ++i) {
    const int32_t k = p.state_keys[i];
    SYMHIP_CHECK(k >= 0 && k < p.n_keys && pos_of[k] < 0, "invalid or repeated optimized key");
    pos_of[k] = i;
    p.tan_off[k] = N;
    off_of_pos[i] = N;
    td_of_pos[i] = tangent_dim(p.types[k]);
    N += td_of_pos[i];
  }
  p.N = N;

  // Adjacency and touched check
  std::vector<char> touched(S, 0);
  p.adj.assign(S, {});
  p.res_off.resize(p.n_factors);
  int64_t M = 0;
  std::vector<int32_t> lin;  // state positions of the factor's optimized linearized args
  for (int32_t f = 0; f < p.n_factors; ++f) {
    const FactorSpec s = factor_spec(p.ftype[f]);
    p.res_off[f] = (int32_t)M;
    M += s.res_dim;
    lin.clear();
    for (int a = 0; a < s.nargs; ++a) {
      if (!(s.lin_mask & (1 << a))) continue;
      const int32_t sp = pos_of[p.fargs[p.fargs_ptr[f] + a]];
      if (sp >= 0) {
        lin.push_back(sp);
        touched[sp] = 1;
      }
    }
    for (size_t i = 0; i < lin.size(); ++i)
      for (size_t j = 0; j < lin.size(); ++j)
        if (lin[i] < lin[j]) p.adj[lin[i]].push_back(lin[j]);
  }
  SYMHIP_CHECK(M < INT32_MAX, "residual dimension exceeds 2^31");
  p.M = (int32_t)M;
  for (int32_t i = 0; i < S; ++i)
    SYMHIP_CHECK(touched[i], "Key is in the state vector but is not optimized by any factor.");
  for (auto& a : p.adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }

  // CSC pattern: column-block of state position k
  p.colptr.assign(N + 1, 0);
  std::vector<int64_t> below(S, 0);  // rows below the diagonal block in each column of block k
  for (int32_t k = 0; k < S; ++k) {
    int64_t b = 0;
    for (int32_t r : p.adj[k]) b += td_of_pos[r];
    below[k] = b;
  }
  {
    int64_t acc = 0;
    for (int32_t k = 0; k < S; ++k)
      for (int32_t l = 0; l < td_of_pos[k]; ++l) {
        p.colptr[off_of_pos[k] + l] = (int32_t)acc;
        acc += (td_of_pos[k] - l) + below[k];
        SYMHIP_CHECK(acc < INT32_MAX, "hessian_lower exceeds 2^31 nonzeros");
      }
    p.colptr[N] = (int32_t)acc;
    p.rowind.resize(acc);
  }
  for (int32_t k = 0; k < S; ++k)
    for (int32_t l = 0; l < td_of_pos[k]; ++l) {
      int32_t q = p.colptr[off_of_pos[k] + l];
      for (int32_t r = l; r < td_of_pos[k]; ++r) p.rowind[q++] = off_of_pos[k] + r;
      for (int32_t rk : p.adj[k])
        for (int32_t r = 0; r < td_of_pos[rk]; ++r) p.rowind[q++] = off_of_pos[rk] + r;
    }
  // position (in rows) of coupled block rk inside the below part of block k
  std::vector<std::vector<int32_t>> adj_pos(S);
  for (int32_t k = 0; k < S; ++k) {
    int32_t acc = 0;
    adj_pos[k].reserve(p.adj[k].size());
    for (int32_t r : p.adj[k]) {
      adj_pos[k].push_back(acc);
      acc += td_of_pos[r];
    }
  }
  auto below_pos = [&](int32_t k, int32_t rk) -> int32_t {
    const auto& a = p.adj[k];
    const auto it = std::lower_bound(a.begin(), a.end(), rk);
    SYMHIP_CHECK(it != a.end() && *it == rk, "internal: coupled key not found");
    return adj_pos[k][it - a.begin()];
  };

  // hessian_storage_col_starts, in the reference's insertion order
  p.cs_ptr.assign(p.n_factors + 1, 0);
  p.col_starts.clear();
  for (int32_t f = 0; f < p.n_factors; ++f) {
    p.cs_ptr[f] = (int32_t)p.col_starts.size();
    const FactorSpec s = factor_spec(p.ftype[f]);
    lin.clear();
    for (int a = 0; a < s.nargs; ++a) {
      if (!(s.lin_mask & (1 << a))) continue;
      const int32_t sp = pos_of[p.fargs[p.fargs_ptr[f] + a]];
      if (sp >= 0) lin.push_back(sp);
    }
    for (size_t ki = 0; ki < lin.size(); ++ki) {
      const int32_t a = lin[ki];
      for (int32_t c = 0; c < td_of_pos[a]; ++c)
        p.col_starts.push_back(p.colptr[off_of_pos[a] + c]);
      for (size_t kj = 0; kj < ki; ++kj) {
        const int32_t b = lin[kj];
        if (off_of_pos[b] < off_of_pos[a]) {
          const int32_t pos = below_pos(b, a);
          for (int32_t c = 0; c < td_of_pos[b]; ++c)
            p.col_starts.push_back(p.colptr[off_of_pos[b] + c] + (td_of_pos[b] - c) + pos);
        } else {
          const int32_t pos = below_pos(a, b);
          for (int32_t c = 0; c < td_of_pos[a]; ++c)
            p.col_starts.push_back(p.colptr[off_of_pos[a] + c] + (td_of_pos[a] - c) + pos);
        }
      }
    }
  }
  p.cs_ptr[p.n_factors] = (int32_t)p.col_starts.size();
}

void build_jacobian_structure(const HostProblem& p, JacobianStructure& j) {
  // rows of column c: the residual rows of every factor whose optimized arguments cover c, in
  // factor order (= row order)
  std::vector<int64_t> count(p.N, 0);
  auto for_columns = [&](int32_t f, auto&& fn) {
    const FactorSpec s = factor_spec(p.ftype[f]);
    int32_t t = 0;
    for (int a = 0; a < s.nargs; ++a) {
      if (!(s.lin_mask & (1 << a))) continue;
      const int32_t k = p.fargs[p.fargs_ptr[f] + a];
      const int32_t td = tangent_dim(p.types[k]);
      for (int32_t c = 0; c < td; ++c) fn(t + c, p.tan_off[k] < 0 ? -1 : p.tan_off[k] + c);
      t += td;
    }
  };
  for (int32_t f = 0; f < p.n_factors; ++f) {
    const int32_t D = factor_spec(p.ftype[f]).res_dim;
    for_columns(f, [&](int32_t, int32_t col) {
      if (col >= 0) count[col] += D;
    });
  }
  j.colptr.assign(p.N + 1, 0);
  for (int32_t c = 0; c < p.N; ++c) {
    j.colptr[c + 1] = (int32_t)(j.colptr[c] + count[c]);
    SYMHIP_CHECK(j.colptr[c + 1] >= j.colptr[c], "jacobian exceeds 2^31 nonzeros");
  }
  j.rowind.assign(j.colptr[p.N], 0);
  std::vector<int64_t> fill(j.colptr.begin(), j.colptr.end() - 1);
  j.sptr.assign(p.n_factors + 1, 0);
  j.starts.clear();
  for (int32_t f = 0; f < p.n_factors; ++f) {
    const FactorSpec s = factor_spec(p.ftype[f]);
    j.sptr[f] = (int32_t)j.starts.size();
    j.starts.resize(j.starts.size() + s.tdim, -1);
    for_columns(f, [&](int32_t t, int32_t col) {
      if (col < 0) return;
      const int64_t q = fill[col];
      j.starts[j.sptr[f] + t] = q;
      for (int32_t r = 0; r < s.res_dim; ++r) j.rowind[q + r] = p.res_off[f] + r;
      fill[col] += s.res_dim;
    });
  }
  j.sptr[p.n_factors] = (int32_t)j.starts.size();
}

}  // namespace symhip
