// Batched per-factor-type linearization kernels (CSC scatter form).
//
// One thread per factor: gather the arguments from the Values storage, evaluate residual and
// analytic Jacobian (geo.h), then scatter-add J^T r into rhs and the lower triangle of J^T J into
// hessian_lower through the reference's per-factor col-start maps
// (opt/linearizer.cc:358-433 gives the loop order reproduced here).
#include "./factor_traits.h"
#include "./linearize.h"
#include "./util.h"

namespace symhip {

namespace {

template <int FTC>
__global__ void __launch_bounds__(128) lin_csc_kernel(int32_t n, const int32_t* __restrict__ arg_off,
                                                      const int32_t* __restrict__ res_off,
                                                      const int32_t* __restrict__ tan_off,
                                                      const int32_t* __restrict__ cs_ptr,
                                                      const int32_t* __restrict__ col_starts,
                                                      const double* __restrict__ values,
                                                      double* __restrict__ residual,
                                                      double* __restrict__ hval,
                                                      double* __restrict__ rhs) {
  using F = FT<FTC>;
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
  if (!hval) return;
  int32_t toff[F::NL];
#pragma unroll
  for (int k = 0; k < F::NL; ++k) toff[k] = tan_off[k * n + f];
  const int32_t* cs = col_starts + cs_ptr[f];
#pragma unroll
  for (int ki = 0; ki < F::NL; ++ki) {
    if (toff[ki] < 0) continue;
    const int tdi = F::td(ki), fi = F::foff(ki);
    for (int c = 0; c < tdi; ++c) {
      double acc = 0;
#pragma unroll
      for (int q = 0; q < F::D; ++q) acc += J[q][fi + c] * r[q];
      unsafeAtomicAdd(&rhs[toff[ki] + c], acc);
    }
    // diagonal block, column by column, lower part
    for (int c = 0; c < tdi; ++c) {
      const int32_t start = *cs++;
      for (int rr = c; rr < tdi; ++rr) {
        double acc = 0;
#pragma unroll
        for (int q = 0; q < F::D; ++q) acc += J[q][fi + rr] * J[q][fi + c];
        unsafeAtomicAdd(&hval[start + rr - c], acc);
      }
    }
    // off-diagonal blocks with earlier optimized linearized args
    for (int kj = 0; kj < ki; ++kj) {
      if (toff[kj] < 0) continue;
      const int tdj = F::td(kj), fj = F::foff(kj);
      if (toff[kj] < toff[ki]) {
        for (int cj = 0; cj < tdj; ++cj) {
          const int32_t start = *cs++;
          for (int rr = 0; rr < tdi; ++rr) {
            double acc = 0;
#pragma unroll
            for (int q = 0; q < F::D; ++q) acc += J[q][fi + rr] * J[q][fj + cj];
            unsafeAtomicAdd(&hval[start + rr], acc);
          }
        }
      } else {
        for (int ci = 0; ci < tdi; ++ci) {
          const int32_t start = *cs++;
          for (int rr = 0; rr < tdj; ++rr) {
            double acc = 0;
#pragma unroll
            for (int q = 0; q < F::D; ++q) acc += J[q][fj + rr] * J[q][fi + ci];
            unsafeAtomicAdd(&hval[start + rr], acc);
          }
        }
      }
    }
  }
}

}  // namespace

void CscLinearizer::setup(const HostProblem& p, hipStream_t s) {
  nnz = p.nnz();
  N = p.N;
  M = p.M;
  groups_.clear();
  std::vector<std::vector<int32_t>> by_type(kNumFactorTypes);
  for (int32_t f = 0; f < p.n_factors; ++f) by_type[p.ftype[f]].push_back(f);
  for (int t = 1; t < kNumFactorTypes; ++t) {
    const auto& fs = by_type[t];
    if (fs.empty()) continue;
    const FactorSpec spec = factor_spec(t);
    const int32_t n = (int32_t)fs.size();
    int nl = 0;
    for (int a = 0; a < spec.nargs; ++a) nl += (spec.lin_mask >> a) & 1;
    std::vector<int32_t> ao((size_t)spec.nargs * n), ro(n), to((size_t)nl * n), cp(n);
    for (int32_t i = 0; i < n; ++i) {
      const int32_t f = fs[i];
      int li = 0;
      for (int a = 0; a < spec.nargs; ++a) {
        const int32_t k = p.fargs[p.fargs_ptr[f] + a];
        ao[(size_t)a * n + i] = p.storage_offset[k];
        if ((spec.lin_mask >> a) & 1) to[(size_t)(li++) * n + i] = p.tan_off[k];
      }
      ro[i] = p.res_off[f];
      cp[i] = p.cs_ptr[f];
    }
    FactorGroup g;
    g.type = t;
    g.n = n;
    g.arg_off.upload(ao, s);
    g.res_off.upload(ro, s);
    g.tan_off.upload(to, s);
    g.cs_ptr.upload(cp, s);
    groups_.push_back(std::move(g));
  }
  col_starts_.upload(p.col_starts, s);
  HIP_OK(hipStreamSynchronize(s));
}

void CscLinearizer::run(const double* values, double* residual, double* hval, double* rhs,
                        hipStream_t s) const {
  if (hval) fill_zero(hval, nnz, s);
  if (rhs) fill_zero(rhs, N, s);
  for (const FactorGroup& g : groups_) {
    const int block = 128;
    const int grid = (g.n + block - 1) / block;
    dispatch_factor_type(g.type, [&]<int C>() {
      hipLaunchKernelGGL(lin_csc_kernel<C>, dim3(grid), dim3(block), 0, s, g.n, g.arg_off.get(),
                         g.res_off.get(), g.tan_off.get(), g.cs_ptr.get(), col_starts_.get(),
                         values, residual, hval, rhs);
    });
    HIP_OK(hipGetLastError());
  }
}

}  // namespace symhip
