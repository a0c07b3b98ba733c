// Batched per-factor-type linearization into the reference's hessian_lower CSC layout, in two
// atomic-free passes (the local-record pass and the gather pass described in linearize.h).
//
// FP64 atomics were the scatter form's cost: at 100 000 Pose3 nodes (115 000 between factors, 90
// hessian and rhs entries each, ~10 M atomic adds per linearization) its launch took 480 us, and
// 81 us with the atomics replaced by plain stores (timing only). The local records (720 B per
// between factor) and the gather move ~0.2 GB instead.
#include "./factor_traits.h"
#include "./linearize.h"
#include "./util.h"

namespace symhip {

namespace {

// Local outputs per factor of type FTC, in the order the pass-1 kernel writes them: for each
// linearized argument i, its rhs entries (td_i), its lower diagonal block column by column
// (td_i (td_i + 1) / 2), then its block with each earlier argument j, rows of i by columns of j
// (td_i td_j).
template <int FTC>
constexpr int local_size() {
  using F = FT<FTC>;
  int L = 0;
  for (int i = 0; i < F::NL; ++i) {
    L += F::td(i) + F::td(i) * (F::td(i) + 1) / 2;
    for (int j = 0; j < i; ++j) L += F::td(i) * F::td(j);
  }
  return L;
}

template <int FTC>
__global__ void __launch_bounds__(128) lin_local_kernel(int32_t n, const int32_t* __restrict__ arg_off,
                                                        const int32_t* __restrict__ res_off,
                                                        const double* __restrict__ values,
                                                        double* __restrict__ residual,
                                                        double* __restrict__ contrib) {
  using F = FT<FTC>;
  constexpr int L = local_size<FTC>();
  const int32_t f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const double* a[F::NA];
#pragma unroll
  for (int k = 0; k < F::NA; ++k) a[k] = values + arg_off[k * n + f];
  double r[F::D];
  double J[F::D][F::T];
  F::template eval<double>(a, r, J);
  if (residual) {
    const int32_t ro = res_off[f];
#pragma unroll
    for (int i = 0; i < F::D; ++i) residual[ro + i] = r[i];
  }
  if (!contrib) return;
  double* out = contrib + (int64_t)f * L;
  int o = 0;
#pragma unroll
  for (int ki = 0; ki < F::NL; ++ki) {
    const int tdi = F::td(ki), fi = F::foff(ki);
#pragma unroll
    for (int c = 0; c < tdi; ++c) {
      double acc = 0;
#pragma unroll
      for (int q = 0; q < F::D; ++q) acc += J[q][fi + c] * r[q];
      out[o++] = acc;
    }
#pragma unroll
    for (int c = 0; c < tdi; ++c)
#pragma unroll
      for (int rr = c; rr < tdi; ++rr) {
        double acc = 0;
#pragma unroll
        for (int q = 0; q < F::D; ++q) acc += J[q][fi + rr] * J[q][fi + c];
        out[o++] = acc;
      }
#pragma unroll
    for (int kj = 0; kj < ki; ++kj) {
      const int tdj = F::td(kj), fj = F::foff(kj);
#pragma unroll
      for (int rr = 0; rr < tdi; ++rr)
#pragma unroll
        for (int cc = 0; cc < tdj; ++cc) {
          double acc = 0;
#pragma unroll
          for (int q = 0; q < F::D; ++q) acc += J[q][fi + rr] * J[q][fj + cc];
          out[o++] = acc;
        }
    }
  }
}

// out[e] = sum of contrib[idx[p]] over p in [ptr[e], ptr[e + 1]), in list order.
__global__ void __launch_bounds__(256) gather_sum_kernel(int64_t n, const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ idx,
                                                         const double* __restrict__ contrib,
                                                         double* __restrict__ out) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  double acc = 0;
  for (int32_t q = ptr[e]; q < ptr[e + 1]; ++q) acc += contrib[idx[q]];
  out[e] = acc;
}

int local_size_of(int type) {
  int L = 0;
  dispatch_factor_type(type, [&]<int C>() { L = local_size<C>(); });
  return L;
}

}  // namespace

void CscLinearizer::setup(const HostProblem& p, hipStream_t s) {
  nnz = p.nnz();
  N = p.N;
  M = p.M;
  groups_.clear();
  std::vector<std::vector<int32_t>> by_type(kNumFactorTypes);
  for (int32_t f = 0; f < p.n_factors; ++f) by_type[p.ftype[f]].push_back(f);
  std::vector<int64_t> base(p.n_factors);  // first contribution entry of each factor
  int64_t total = 0;
  for (int t = 1; t < kNumFactorTypes; ++t) {
    const auto& fs = by_type[t];
    if (fs.empty()) continue;
    const FactorSpec spec = factor_spec(t);
    const int32_t n = (int32_t)fs.size();
    const int L = local_size_of(t);
    std::vector<int32_t> ao((size_t)spec.nargs * n), ro(n);
    for (int32_t i = 0; i < n; ++i) {
      const int32_t f = fs[i];
      for (int a = 0; a < spec.nargs; ++a)
        ao[(size_t)a * n + i] = p.storage_offset[p.fargs[p.fargs_ptr[f] + a]];
      ro[i] = p.res_off[f];
      base[f] = total + (int64_t)i * L;
    }
    FactorGroup g;
    g.type = t;
    g.n = n;
    g.arg_off.upload(ao, s);
    g.res_off.upload(ro, s);
    g.contrib_off = total;
    groups_.push_back(std::move(g));
    total += (int64_t)n * L;
  }
  SYMHIP_CHECK(total < INT32_MAX, "linearization exceeds 2^31 local contributions");
  contrib_.resize(std::max<int64_t>(1, total));
  // Gather lists, replaying each factor's local layout against the reference's storage map
  // (col_starts in the order problem.cc builds it); factors in their given order.
  auto replay = [&](auto&& emit_h, auto&& emit_r) {
    for (int32_t f = 0; f < p.n_factors; ++f) {
      const FactorSpec spec = factor_spec(p.ftype[f]);
      int td[6], toff[6], nl = 0;
      for (int a = 0; a < spec.nargs; ++a) {
        if (!((spec.lin_mask >> a) & 1)) continue;
        const int32_t k = p.fargs[p.fargs_ptr[f] + a];
        td[nl] = tangent_dim(p.types[k]);
        toff[nl++] = p.tan_off[k];
      }
      const int32_t* cs = p.col_starts.data() + p.cs_ptr[f];
      int64_t o = base[f];
      for (int ki = 0; ki < nl; ++ki) {
        const int tdi = td[ki];
        if (toff[ki] < 0) {
          o += tdi + tdi * (tdi + 1) / 2;
          for (int kj = 0; kj < ki; ++kj) o += tdi * td[kj];
          continue;
        }
        for (int c = 0; c < tdi; ++c) emit_r(toff[ki] + c, o + c);
        o += tdi;
        for (int c = 0; c < tdi; ++c) {
          const int32_t start = *cs++;
          for (int rr = c; rr < tdi; ++rr) emit_h(start + rr - c, o++);
        }
        for (int kj = 0; kj < ki; ++kj) {
          const int tdj = td[kj];
          if (toff[kj] >= 0) {
            if (toff[kj] < toff[ki]) {
              // stored as columns of argument j, rows of argument i
              for (int cj = 0; cj < tdj; ++cj) {
                const int32_t start = *cs++;
                for (int rr = 0; rr < tdi; ++rr) emit_h(start + rr, o + rr * tdj + cj);
              }
            } else {
              // stored as columns of argument i, rows of argument j
              for (int ci = 0; ci < tdi; ++ci) {
                const int32_t start = *cs++;
                for (int rr = 0; rr < tdj; ++rr) emit_h(start + rr, o + ci * tdj + rr);
              }
            }
          }
          o += tdi * tdj;
        }
      }
    }
  };
  std::vector<int32_t> hptr(nnz + 1, 0), rptr(N + 1, 0);
  replay([&](int64_t e, int64_t) { ++hptr[e + 1]; }, [&](int64_t e, int64_t) { ++rptr[e + 1]; });
  for (int64_t e = 0; e < nnz; ++e) hptr[e + 1] += hptr[e];
  for (int32_t e = 0; e < N; ++e) rptr[e + 1] += rptr[e];
  std::vector<int32_t> hidx(std::max<int64_t>(1, hptr[nnz])), ridx(std::max(1, rptr[N]));
  {
    std::vector<int32_t> hf(hptr.begin(), hptr.end() - 1), rf(rptr.begin(), rptr.end() - 1);
    replay([&](int64_t e, int64_t q) { hidx[hf[e]++] = (int32_t)q; },
           [&](int64_t e, int64_t q) { ridx[rf[e]++] = (int32_t)q; });
  }
  hptr_.upload(hptr, s);
  hidx_.upload(hidx, s);
  rptr_.upload(rptr, s);
  ridx_.upload(ridx, s);
  HIP_OK(hipStreamSynchronize(s));
}

void CscLinearizer::run(const double* values, double* residual, double* hval, double* rhs,
                        hipStream_t s) const {
  const bool sums = hval || rhs;
  for (const FactorGroup& g : groups_) {
    const int block = 128;
    const int grid = (g.n + block - 1) / block;
    dispatch_factor_type(g.type, [&]<int C>() {
      hipLaunchKernelGGL(lin_local_kernel<C>, dim3(grid), dim3(block), 0, s, g.n,
                         g.arg_off.get(), g.res_off.get(), values, residual,
                         sums ? contrib_.get() + g.contrib_off : static_cast<double*>(nullptr));
    });
    HIP_OK(hipGetLastError());
  }
  if (hval)
    hipLaunchKernelGGL(gather_sum_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz,
                       hptr_.get(), hidx_.get(), contrib_.get(), hval);
  if (rhs)
    hipLaunchKernelGGL(gather_sum_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, (int64_t)N,
                       rptr_.get(), ridx_.get(), contrib_.get(), rhs);
}

}  // namespace symhip
