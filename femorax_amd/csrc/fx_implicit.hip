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
// (NormUnbalance ||B||, NormDispIncr ||dU|| or EnergyIncr 0.5 |dU . B|, :239-246). A converged
// attempt commits the material history (:365-372); a failed one restores its start, and the
// step is retried and substepped as the reference's recover does (:385-402).
#include <hip/hip_runtime.h>

#include <cmath>
#include <functional>

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

__global__ void axpy1_kernel(int32_t n, const double* __restrict__ du, double* __restrict__ u) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) u[q] += du[q];
}

__global__ void mul_kernel(int32_t n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ z) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) z[q] = x[q] * y[q];
}

}  // namespace

int multiply64(fx_ctx* ctx, int32_t n, const double* x, const double* y, double* z) {
  if (n > 0) hipLaunchKernelGGL(mul_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, ctx->stream, n, x, y, z);
  FX_HIP(ctx, hipGetLastError());
  return FX_OK;
}

namespace {

__global__ void newmark_residual_kernel(int32_t n, const double* f, const double* ma, const double* mr, double ag,
                                        double* b) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double r = -f[q];
  if (ma) r -= ma[q];
  if (mr) r -= ag * mr[q];
  b[q] = r;
}

}  // namespace

// PathTimeSeries::getFactor (ground_motion.cpp:43-74): the interval [time[l], time[l + 1]] found
// by walking on from the previous call's l, forward or back (loc keeps it); a time equal to the
// walk's lower end gives that sample exactly; 0 outside the data.
double path_time(const fx_time_series& s, int32_t& loc, double t) {
  const double* v = s.values;
  const double* tm = s.time;
  const int32_t n = s.count;
  int32_t l = loc;
  double t1 = tm[l];
  if (t < t1 && l == 0) return 0.0;
  if (t == t1) return s.factor * v[l];
  if (t > t1 && l == n - 1) return 0.0;
  double t2 = tm[l + 1];
  if (t > t2) {
    while (t > t2 && l < n - 2) {
      ++l;
      t1 = t2;
      t2 = tm[l + 1];
    }
    loc = l;
    if (t > t2) return 0.0;
  } else if (t < t1) {
    while (t < t1 && l > 0) {
      --l;
      t2 = t1;
      t1 = tm[l];
    }
    loc = l;
    if (t < t1) return 0.0;
  }
  return s.factor * (v[l] + (v[l + 1] - v[l]) * (t - t1) / (t2 - t1));
}

// The series' value times the pattern factor: PathSeries::getFactor for a uniform series
// (ground_motion.cpp:17-33), PathTimeSeries for a time array; 0 before time 0
// (GroundMotions::evaluate, :121-145).
double series_value(const fx_time_series& s, double fact, double t, int32_t& loc) {
  if (t < 0.0) return 0.0;
  if (s.time) return fact * path_time(s, loc, t);
  double a = 0.0;
  if (!(t < s.start)) {
    const double incr = (t - s.start) / s.dt;
    const long long incr1 = static_cast<long long>(std::floor(incr));
    const long long incr2 = incr1 + 1;
    if (incr2 >= s.count) {
      a = s.use_last ? s.factor * s.values[s.count - 1] : 0.0;
    } else {
      const double value1 = s.values[incr1], value2 = s.values[incr2];
      a = s.factor * (value1 + (value2 - value1) * (incr - static_cast<double>(incr1)));
    }
  }
  return fact * a;
}

int newmark_run(fx_batch* b, const fx_newmark_desc& d, double* u, double* v, double* a, double* du,
                fx_newmark_result& out) {
  fx_ctx* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const int32_t n = b->neq;
  out = fx_newmark_result{};
  out.time = d.time;
  // a time-array series walks on from the caller's interval (or interval 0) and keeps its walk
  // from attempt to attempt
  int32_t loc = d.ground_series && d.ground_series->loc ? *d.ground_series->loc : 0;
  if (n == 0 || d.nsteps <= 0) return FX_OK;
  const size_t N = static_cast<size_t>(n);
  double* W = b->dyn;
  double *U = u, *V = v, *A = a, *dU = du;
  double *Ut = W + 4 * N, *Vt = W + 5 * N, *At = W + 6 * N, *F = W + 7 * N, *MA = W + 8 * N, *B = W + 9 * N,
         *MR = W + 10 * N, *SCR = W + 11 * N, *dU0 = W + 12 * N;
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
  if (ground) {
    // mass_times_ground (internal_force.cpp:601-626): M r with r the rigid unit translation at
    // every element DOF of the direction, fixed ones included. The element mass (consistent or
    // lumped) times a rigid translation is its row sums rho sum_gp N_j dvol, so M r is the lumped
    // mass at the direction's equations.
    std::vector<double> r(N, 0.0);
    for (int32_t q = 0; q < n; ++q)
      if (b->eq_dir[q] == d.ground_dir) r[b->internal ? b->internal_of[q] : q] = 1.0;
    FX_HIP(ctx, hipMemcpyAsync(SCR, r.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    FX_HIP(ctx, hipStreamSynchronize(st));
    hipLaunchKernelGGL(mul_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, b->lumped_m, SCR, MR);
    FX_HIP(ctx, hipGetLastError());
  }
  // node recorders after every commit (implicit_stage.cpp:371-381)
  const bool rec = recorder_active(d.recorder);
  const int32_t* rmap = nullptr;
  if (rec)
    if (int rc = batch_internal_map(b, &rmap)) return rc;
  const double gamma = d.gamma, beta = d.beta;
  const bool is_static = d.analysis == 1;
  const bool linear = d.algorithm == 1, modified = d.algorithm == 2;
  const bool factor_once = modified && d.factor_once, initial = modified && d.modified_initial;
  fx_solve_options so{d.cg_rtol, 0.0, d.cg_max_it};
  double ag = 0.0;
  // the operator's coefficients per step size (:146-150): static K (or K_0), transient
  // K + gamma/(beta h) C + 1/(beta h^2) M
  auto coefficients = [&](double h) -> fx_coeffs {
    if (is_static) return fx_coeffs{1.0, 0.0, 0.0, initial ? 1 : 0};
    return fx_coeffs{1.0, gamma / (beta * h), 1.0 / (beta * h * h), initial ? 1 : 0};
  };
  // form_unbalance (:212-219): B = -f(u, v) - M a - a_g M r; static stages: the elements see v = 0,
  // no inertia
  auto unbalance = [&]() -> int {
    if (int rc = force64(b, U, is_static ? nullptr : V, F, SCR)) return rc;
    if (!is_static)
      if (int rc = operator_apply64(b, mass, A, MA)) return rc;
    hipLaunchKernelGGL(newmark_residual_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, F,
                       is_static ? nullptr : MA, ground ? MR : nullptr, ag, B);
    FX_HIP(ctx, hipGetLastError());
    return FX_OK;
  };
  // test_value (:239-246): ||B||, ||dU|| or 0.5 |dU . B|, dots in the VectorOps order
  auto test_value = [&](double& val) -> int {
    double s = 0.0;
    if (d.test == 2) {
      if (int rc = dot64(ctx, n, dU, B, &s)) return rc;
      val = 0.5 * std::fabs(s);
      return FX_OK;
    }
    const double* x = d.test == 0 ? B : dU;
    if (int rc = dot64(ctx, n, x, x, &s)) return rc;
    val = std::sqrt(s);
    return FX_OK;
  };
  // update_tangent (:173-193): the tangents at u (stored tangents of history materials); the
  // Jacobi diagonal is formed by each solve from the current tangents
  double tangent_h = d.dt;
  bool kept_tangent = false;
  auto update_tangent = [&](double h) -> int {
    tangent_h = h;
    if (b->pimy)
      if (int rc = pimy_force64(b, U, nullptr, SCR, true)) return rc;
    return FX_OK;
  };
  // The solver KrylovSolver::setup picks (krylov.cpp:10-22): CG for an SPD operator (with the
  // GMRES fallback when the workspace is there), GMRES otherwise; or the mixed-precision CG.
  auto solve = [&](const fx_coeffs& k, fx_solve_result& r) -> int {
    if (d.mixed) return cg_mixed_solver(b, k, so, d.inner_rtol, B, dU, r);
    if (b->gm) return krylov64_solver(b, k, so, B, dU, r);
    return cg64_solver(b, k, so, B, dU, r);
  };
  // LoadControl (:262-269): the increment, its bounds, numIter and the last attempt's iterations
  double dlambda = d.load_increment;
  const bool adapt = d.load_min_increment != 0.0 && d.load_max_increment != 0.0;
  const double dmin = adapt ? d.load_min_increment : d.load_increment;
  const double dmax = adapt ? d.load_max_increment : d.load_increment;
  const double spec_iter = adapt ? static_cast<double>(d.load_num_iter) : 1.0;
  double last_iter = spec_iter;
  int32_t step = 0;  // the step of size dt being recovered (its ground_accel entry)
  // One attempt of size h (:274-372): predictor (transient) or the load increment (static),
  // unbalance, the algorithm; a failure restores u (and v, a), the initial guess and the time.
  auto attempt = [&](double h, bool& ok) -> int {
    const double time0 = out.time;
    if (is_static) {
      dlambda *= spec_iter / last_iter;
      if (dlambda < dmin) dlambda = dmin;
      else if (dlambda > dmax) dlambda = dmax;
      if (int rc = copy64(ctx, n, U, Ut)) return rc;
      out.time += dlambda;
      last_iter = 0;
    } else {
      const double a1 = 1.0 - gamma / beta, a2 = h * (1.0 - 0.5 * gamma / beta);
      const double a3 = -1.0 / (beta * h), a4 = 1.0 - 0.5 / beta;
      hipLaunchKernelGGL(newmark_predict_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, U, V, A, Ut, Vt, At,
                         a1, a2, a3, a4);
      out.time += h;
    }
    if (int rc = copy64(ctx, n, dU, dU0)) return rc;  // the solver's initial guess, restored on failure
    ag = !ground ? 0.0 : d.ground_series ? series_value(*d.ground_series, d.ground_factor, out.time, loc) : d.ground_accel[step];
    const fx_coeffs k = coefficients(h);
    auto update = [&]() -> int {
      if (is_static)
        hipLaunchKernelGGL(axpy1_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, dU, U);
      else
        hipLaunchKernelGGL(newmark_update_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, st, n, dU, U, V, A,
                           k.damping, k.mass);
      FX_HIP(ctx, hipGetLastError());
      return FX_OK;
    };
    int iterations = 0;
    ok = true;
    if (linear) {  // :301-313: the tangent, the unbalance, one solve that must converge
      if (int rc = update_tangent(h)) return rc;
      if (int rc = unbalance()) return rc;
      fx_solve_result r;
      if (int rc = solve(k, r)) return rc;
      out.krylov_iterations += r.iterations;
      if (!r.converged) ok = false;
      else if (int rc = update()) return rc;
      iterations = 1;
    } else {
      if (int rc = unbalance()) return rc;
      // ModifiedNewton: the tangent once per attempt, after the first unbalance; -factoronce: once
      // while it stays valid (:324-328)
      if (modified && !(factor_once && kept_tangent && h == tangent_h))
        if (int rc = update_tangent(h)) return rc;
      kept_tangent = factor_once;
      int iter = 1;
      for (;;) {
        if (!modified)
          if (int rc = update_tangent(h)) return rc;
        fx_solve_result r;
        if (int rc = solve(k, r)) return rc;
        out.krylov_iterations += r.iterations;
        if (!r.converged && !(r.final_residual <= 1e-6 * r.rhs_norm)) {
          ok = false;
          break;
        }
        if (int rc = update()) return rc;
        if (int rc = unbalance()) return rc;
        double value = 0.0;
        if (int rc = test_value(value)) return rc;
        iterations = iter;
        out.last_test = value;
        if (value <= d.tol) break;
        if (iter >= d.max_iter) {
          ok = false;
          kept_tangent = false;  // OpenSees: factorOnce = 1 after a failed test
          break;
        }
        ++iter;
      }
    }
    out.newton_iterations += iterations;
    if (is_static) last_iter = iterations;  // LoadControl counts the failed attempt's updates too
    if (!ok) {
      if (int rc = copy64(ctx, n, Ut, U)) return rc;
      if (!is_static) {
        if (int rc = copy64(ctx, n, Vt, V)) return rc;
        if (int rc = copy64(ctx, n, At, A)) return rc;
      }
      if (int rc = copy64(ctx, n, dU0, dU)) return rc;
      out.time = time0;
      ++out.failed_attempts;
      return FX_OK;
    }
    out.max_newton_iterations = std::max(out.max_newton_iterations, iterations);
    ++out.commits;
    if (b->pimy) std::swap(b->state_c, b->state_t);  // commit_state
    if (rec)  // static stages record the committed velocity, which they keep
      if (int rc = recorder_record(d.recorder, out.time, n, U, V, A, rmap)) return rc;
    return FX_OK;
  };
  // recover (:385-402): retries at the same size, then substeps one level down
  std::function<int(double, int, bool&)> recover = [&](double h, int level, bool& ok) -> int {
    if (int rc = attempt(h, ok)) return rc;
    if (ok) return FX_OK;
    for (int r = 1; r <= d.max_retries; ++r) {
      ++out.retries;
      if (int rc = attempt(h, ok)) return rc;
      if (ok) return FX_OK;
    }
    if (is_static || level >= d.num_sublevels) return FX_OK;  // static stages retry only (:262-264)
    ++out.substepped;
    const double hs = h / static_cast<double>(d.num_substeps);
    for (int i = 0; i < d.num_substeps; ++i) {
      if (int rc = recover(hs, level + 1, ok)) return rc;
      if (!ok) return FX_OK;
    }
    return FX_OK;
  };
  for (step = 0; step < d.nsteps; ++step) {
    bool ok = false;
    if (int rc = recover(d.dt, 0, ok)) return rc;
    if (!ok) {
      out.failed = 1;
      break;
    }
    out.steps = step + 1;
  }
  if (d.ground_series && d.ground_series->loc) *d.ground_series->loc = loc;
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
