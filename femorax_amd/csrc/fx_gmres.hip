// Flexible GMRES(m) with classical Gram-Schmidt run twice (CGS2) and the Jacobi
// preconditioner on the right (src/solvers/gmres.cpp:32-123), and the Krylov driver that runs
// CG and falls back to GMRES from the same initial guess when CG meets p^T A p <= 0
// (src/solvers/krylov.cpp:25-42).
//
// The basis V (m + 1 vectors) and the preconditioned directions Z (m vectors) live in the
// batch's GMRES workspace (FX_BATCH_GMRES); the Hessenberg matrix, rotations and the small
// triangular solve stay on the host. Each projection coefficient is a dot in the VectorOps
// order (BlockOps::dots, block_ops.cpp:59-78); the corrections add sum_c coef_c V(:, c) per
// row (BlockOps::combine :80-92).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"

namespace fx {

namespace {

constexpr int kThreads = 256;
inline dim3 grid(int64_t n) { return dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)); }

__global__ void gm_residual_kernel(int64_t n, const double* __restrict__ b, const double* __restrict__ w,
                                   double* __restrict__ r) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] + -1.0 * w[i];
}
// y = 0 + s x (VectorOps::fill then axpy)
__global__ void gm_scaled_kernel(int64_t n, double s, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = 0.0 + s * x[i];
}
__global__ void gm_precond_kernel(int64_t n, const double* __restrict__ inv, const double* __restrict__ v,
                                  double* __restrict__ z) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) z[i] = inv[i] * v[i];
}
// w(i) += sum_c coef(c) V(i, c), V column-major with stride n
__global__ void gm_combine_kernel(int64_t n, int k, const double* __restrict__ coef, const double* __restrict__ V,
                                  double* __restrict__ w) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < k; ++c) s += coef[c] * V[c * n + i];
  w[i] += s;
}

}  // namespace

int gmres64_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* bi, double* xi,
                   fx_solve_result& res) {
  fx_ctx* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = b->neq;
  const int m = kGmresRestart;
  if (!b->gm) {
    ctx->error = "gmres: the batch was created without FX_BATCH_GMRES";
    return FX_EUNSUP;
  }
  double *w = b->work, *r = b->work + n, *inv = b->work + 4 * n, *d = b->work + 5 * n;
  double *V = b->gm, *Z = b->gm + (m + 1) * n, *coef = b->gm + (2 * m + 2) * n;
  double* dscratch = b->gm + (2 * m + 3) * n + m + 1;  // the block dots' partials
  if (int rc = jacobi_inverse_public(b, k, d, inv)) return rc;
  res = fx_solve_result{};
  auto dot = [&](const double* x, const double* y, double& out) { return dot64(ctx, n, x, y, &out); };
  double bb = 0.0;
  if (int rc = dot(bi, bi, bb)) return rc;
  res.rhs_norm = std::sqrt(bb);
  const double tol = std::max(o.relative_tolerance * res.rhs_norm, o.absolute_tolerance);
  std::vector<double> H(static_cast<size_t>((m + 1) * m), 0.0), cs(m), sn(m), g(m + 1), h(m + 1), h2(m + 1), y(m + 1);
  auto Hij = [&](int i, int j) -> double& { return H[static_cast<size_t>(j) * (m + 1) + i]; };
  auto combine = [&](const double* base, int kk, const std::vector<double>& c, double a, double* out) -> int {
    std::vector<double> cc(kk);
    for (int i = 0; i < kk; ++i) cc[i] = a * c[i];
    FX_HIP(ctx, hipMemcpyAsync(coef, cc.data(), sizeof(double) * kk, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(gm_combine_kernel, grid(n), dim3(kThreads), 0, st, n, kk, coef, base, out);
    FX_HIP(ctx, hipGetLastError());
    return FX_OK;
  };
  bool first = true;
  for (;;) {
    if (int rc = operator_apply64(b, k, xi, w)) return rc;
    hipLaunchKernelGGL(gm_residual_kernel, grid(n), dim3(kThreads), 0, st, n, bi, w, r);
    double rr = 0.0;
    if (int rc = dot(r, r, rr)) return rc;
    const double beta = std::sqrt(rr);
    if (first) res.initial_residual = beta;
    first = false;
    res.final_residual = beta;
    if (beta <= tol) {
      res.converged = 1;
      return FX_OK;
    }
    if (res.iterations >= o.max_iterations) return FX_OK;
    hipLaunchKernelGGL(gm_scaled_kernel, grid(n), dim3(kThreads), 0, st, n, 1.0 / beta, r, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m && res.iterations < o.max_iterations; ++j) {
      double *vj = V + static_cast<size_t>(j) * n, *zj = Z + static_cast<size_t>(j) * n;
      hipLaunchKernelGGL(gm_precond_kernel, grid(n), dim3(kThreads), 0, st, n, inv, vj, zj);
      if (int rc = operator_apply64(b, k, zj, w)) return rc;
      for (int pass = 0; pass < 2; ++pass) {  // CGS2: one block of dots, one block update per pass
        std::vector<double>& hh = pass == 0 ? h : h2;
        if (int rc = multidot64(ctx, n, V, j + 1, w, dscratch, hh.data())) return rc;
        if (int rc = combine(V, j + 1, hh, -1.0, w)) return rc;
      }
      for (int i = 0; i <= j; ++i) Hij(i, j) = h[i] + h2[i];
      double hn2 = 0.0;
      if (int rc = dot(w, w, hn2)) return rc;
      const double hn = std::sqrt(hn2);
      Hij(j + 1, j) = hn;
      if (hn > 0.0)
        hipLaunchKernelGGL(gm_scaled_kernel, grid(n), dim3(kThreads), 0, st, n, 1.0 / hn, w,
                           V + static_cast<size_t>(j + 1) * n);
      for (int i = 0; i < j; ++i) {
        const double a = Hij(i, j), c = Hij(i + 1, j);
        Hij(i, j) = cs[i] * a + sn[i] * c;
        Hij(i + 1, j) = -sn[i] * a + cs[i] * c;
      }
      const double a = Hij(j, j), c = Hij(j + 1, j);
      const double dd = std::hypot(a, c);
      cs[j] = dd > 0.0 ? a / dd : 1.0;
      sn[j] = dd > 0.0 ? c / dd : 0.0;
      Hij(j, j) = dd;
      Hij(j + 1, j) = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++res.iterations;
      res.final_residual = std::fabs(g[j + 1]);
      if (res.final_residual <= tol || hn == 0.0) {
        ++j;
        break;
      }
    }
    for (int i = j - 1; i >= 0; --i) {
      double s = g[i];
      for (int kk = i + 1; kk < j; ++kk) s -= Hij(i, kk) * y[kk];
      y[i] = s / Hij(i, i);
    }
    if (int rc = combine(Z, j, y, 1.0, xi)) return rc;
    if (res.final_residual <= tol) {
      if (int rc = operator_apply64(b, k, xi, w)) return rc;
      hipLaunchKernelGGL(gm_residual_kernel, grid(n), dim3(kThreads), 0, st, n, bi, w, r);
      double r2 = 0.0;
      if (int rc = dot(r, r, r2)) return rc;
      res.final_residual = std::sqrt(r2);
      if (res.final_residual <= tol) {
        res.converged = 1;
        return FX_OK;
      }
    }
    if (res.iterations >= o.max_iterations) return FX_OK;
  }
}

int krylov64_solver(fx_batch* b, const fx_coeffs& k, const fx_solve_options& o, const double* bi, double* xi,
                    fx_solve_result& res) {
  const int64_t n = b->neq;
  if (!operator_spd(b, k)) {  // KrylovSolver::setup (krylov.cpp:10-22): GMRES from the start
    if (!b->gm) {
      b->ctx->error = "Krylov solve: the operator is not symmetric positive definite, so it needs GMRES (create "
                      "the batch with FX_BATCH_GMRES)";
      return FX_EINVAL;
    }
    res = fx_solve_result{};
    return gmres64_solver(b, k, o, bi, xi, res);
  }
  double* x0 = b->gm ? b->gm + (2 * kGmresRestart + 2) * n + kGmresRestart + 1 : nullptr;
  if (x0)
    if (int rc = copy64(b->ctx, n, xi, x0)) return rc;
  if (int rc = cg64_solver(b, k, o, bi, xi, res)) return rc;
  if (!res.indefinite || !x0) return FX_OK;
  const int cg_iterations = res.iterations;
  if (int rc = copy64(b->ctx, n, x0, xi)) return rc;
  if (int rc = gmres64_solver(b, k, o, bi, xi, res)) return rc;
  res.iterations += cg_iterations;
  res.indefinite = 1;
  return FX_OK;
}

}  // namespace fx
