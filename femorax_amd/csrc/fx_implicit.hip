// The implicit Newmark stage with Newton iterations (src/exec/implicit_stage.cpp): one call runs
// nsteps steps on the device, the host only reading the scalars each decision needs (the
// CG stop test, the Newton test value).
//
// Per step: the predictor (newmark_predict :29-40, constants :283-286), the unbalance
// B = -f(u, v) - M a - a_g M r (form_unbalance :212-219, form_residual :55-68), then Newton:
// the tangents at u (update_tangent :180-193: stored tangents of history materials, Jacobi set
// up on the current operator), PCG on K_eff = K + gamma/(beta h) C + 1/(beta h^2) M with the
// previous increment as initial guess (an increment whose solve reduced the residual by 1e-6 is
// accepted, :223-238), the update (newmark_update :43-52), a new unbalance and the test value
// (NormUnbalance ||B|| or NormDispIncr ||dU||, :239-246). A converged step commits the material
// history (:365-372); a failed one restores the step's start and ends the call.
#include <hip/hip_runtime.h>

#include <cmath>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"

namespace fx {

namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

__global__ void newmark_predict_kernel(int32_t n, double* u, double* v, double* a, double* ut, double* vt, double* at,
                                       double a1, double a2, double a3, double a4) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  ut[q] = u[q];
  vt[q] = v[q];
  at[q] = a[q];
  v[q] = v[q] * a1 + at[q] * a2;
  a[q] = a[q] * a4 + vt[q] * a3;
}

__global__ void newmark_update_kernel(int32_t n, const double* du, double* u, double* v, double* a, double c2,
                                      double c3) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  u[q] += du[q];
  v[q] += du[q] * c2;
  a[q] += du[q] * c3;
}

__global__ void newmark_residual_kernel(int32_t n, const double* f, const double* ma, const double* mr, double ag,
                                        double* b) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double r = -f[q];
  r -= ma[q];
  if (mr) r -= ag * mr[q];
  b[q] = r;
}

}  // namespace

int newmark_run(fx_batch* b, const fx_newmark_desc& d, double* u, double* v, double* a, double* du,
                fx_newmark_result& out) {
  fx_ctx* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const int32_t n = b->neq;
  out = fx_newmark_result{};
  if (n == 0 || d.nsteps <= 0) return FX_OK;
  const size_t N = static_cast<size_t>(n);
  double* W = b->dyn;
  double *U = u, *V = v, *A = a, *dU = du;
  double *Ut = W + 4 * N, *Vt = W + 5 * N, *At = W + 6 * N, *F = W + 7 * N, *MA = W + 8 * N, *B = W + 9 * N,
         *MR = W + 10 * N, *SCR = W + 11 * N;
  if (b->internal) {
    U = W;
    V = W + N;
    A = W + 2 * N;
    dU = W + 3 * N;
    if (int rc = to_internal<double>(b, u, U)) return rc;
    if (int rc = to_internal<double>(b, v, V)) return rc;
    if (int rc = to_internal<double>(b, a, A)) return rc;
    if (int rc = to_internal<double>(b, du, dU)) return rc;
  }
  const fx_coeffs mass{0.0, 0.0, 1.0, 0};
  const bool ground = d.ground_dir >= 0;
  if (ground) {  // mass_times_ground: M r, r = 1 at the equations of the excitation's direction
    std::vector<double> r(N, 0.0);
    for (int32_t q = 0; q < n; ++q)
      if (b->eq_dir[q] == d.ground_dir) r[b->internal ? b->internal_of[q] : q] = 1.0;
    FX_HIP(ctx, hipMemcpyAsync(SCR, r.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    FX_HIP(ctx, hipStreamSynchronize(st));
    if (int rc = operator_apply64(b, mass, SCR, MR)) return rc;
  }
  const double gamma = d.gamma, beta = d.beta, h = d.dt;
  const fx_coeffs k{1.0, gamma / (beta * h), 1.0 / (beta * h * h), 0};
  const double a1 = 1.0 - gamma / beta, a2 = h * (1.0 - 0.5 * gamma / beta);
  const double a3 = -1.0 / (beta * h), a4 = 1.0 - 0.5 / beta;
  fx_solve_options so{d.cg_rtol, 0.0, d.cg_max_it};
  auto unbalance = [&](double ag) -> int {
    if (int rc = force64(b, U, V, F, SCR)) return rc;
    if (int rc = operator_apply64(b, mass, A, MA)) return rc;
    hipLaunchKernelGGL(newmark_residual_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, F, MA,
                       ground ? MR : nullptr, ag, B);
    FX_HIP(ctx, hipGetLastError());
    return FX_OK;
  };
  auto norm = [&](const double* x, double& val) -> int {
    double s = 0.0;
    if (int rc = dot64(ctx, n, x, x, &s)) return rc;
    val = std::sqrt(s);
    return FX_OK;
  };
  for (int32_t step = 0; step < d.nsteps; ++step) {
    hipLaunchKernelGGL(newmark_predict_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, U, V, A, Ut, Vt, At, a1,
                       a2, a3, a4);
    const double ag = ground && d.ground_accel ? d.ground_accel[step] : 0.0;
    if (int rc = unbalance(ag)) return rc;
    int iter = 1;
    bool ok = true;
    for (;;) {
      if (b->pimy)  // update_tangent: the tangents at u
        if (int rc = pimy_force64(b, U, SCR, true)) return rc;
      fx_solve_result r;
      if (int rc = b->gm ? krylov64_solver(b, k, so, B, dU, r) : cg64_solver(b, k, so, B, dU, r)) return rc;
      out.krylov_iterations += r.iterations;
      if (!r.converged && !(r.final_residual <= 1e-6 * r.rhs_norm)) {
        ok = false;
        break;
      }
      hipLaunchKernelGGL(newmark_update_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, dU, U, V, A, k.damping,
                         k.mass);
      if (int rc = unbalance(ag)) return rc;
      double value = 0.0;
      if (int rc = norm(d.test == 0 ? B : dU, value)) return rc;
      out.last_test = value;
      if (value <= d.tol) break;
      if (iter >= d.max_iter) {
        ok = false;
        break;
      }
      ++iter;
    }
    out.newton_iterations += iter;
    out.max_newton_iterations = std::max(out.max_newton_iterations, iter);
    if (!ok) {
      if (int rc = copy64(ctx, n, Ut, U)) return rc;
      if (int rc = copy64(ctx, n, Vt, V)) return rc;
      if (int rc = copy64(ctx, n, At, A)) return rc;
      out.failed = 1;
      break;
    }
    if (b->pimy) std::swap(b->state_c, b->state_t);  // commit_state
    out.steps = step + 1;
  }
  if (b->internal) {
    if (int rc = to_external<double>(b, U, u)) return rc;
    if (int rc = to_external<double>(b, V, v)) return rc;
    if (int rc = to_external<double>(b, A, a)) return rc;
    if (int rc = to_external<double>(b, dU, du)) return rc;
  }
  FX_HIP(ctx, hipStreamSynchronize(st));
  return FX_OK;
}

}  // namespace fx
