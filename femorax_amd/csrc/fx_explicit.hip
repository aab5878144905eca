// Internal force of elastic stdBrick batches and the explicit CentralDifference and
// ExplicitDifference loops.
//
// For ElasticIsotropic the element force is linear: f = K u + C v - F_body (std_brick.hpp:
// 493-543 with element_operator's Factors), so the force runs on the operator kernels (K u,
// and C v with the operator coefficients (0, 1, 0)) plus a combine kernel; F_body is a vector
// built once per batch. The CentralDifference step (explicit_analysis.cpp:87-117, 712-722)
// is one device-resident loop: per step the garbage velocity v_e = -c2 u(n-1) (only when the
// elements damp), the force, the point loads, and one fused update per equation that forms
// R = -f - M a_g + P, solves the diagonal system (c3 M + c2 C) U = R and writes U, the recorded
// velocity and acceleration. U goes into u(n-1)'s buffer and the two buffers swap roles, so the
// update writes 3 vectors, not 4. ExplicitDifference (explicit_analysis.cpp:60-85) runs its
// predictor, the force at the predicted velocity and one fused update (a = R / m, the recorded
// velocity). Both add the ground-motion load of a uniform excitation (:48-58).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"

namespace fx {

namespace {

constexpr int kThreads = 256;

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

__global__ void scale_kernel(int32_t n, double s, const double* __restrict__ x, double* __restrict__ y) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) y[q] = x[q] * s;
}

// f = Ku (+ Cv) - F_body
__global__ void force_combine_kernel(int32_t n, const double* __restrict__ ku, const double* __restrict__ cv,
                                     const double* __restrict__ body, double* __restrict__ f) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double v = ku[q];
  if (cv) v += cv[q];
  if (body) v -= body[q];
  f[q] = v;
}

struct PointLoads {
  int32_t n;
  int32_t eq[kMaxPointLoads];
  double val[kMaxPointLoads];
};

// f(q) -= P(q): the point loads join the residual R = -f + P
__global__ void point_load_kernel(PointLoads p, double* f) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    for (int j = 0; j < p.n; ++j) f[p.eq[j]] -= p.val[j];
}

struct CdArgs {
  int32_t n;
  double c2, c3, dt;
  const double* f;     // K u (+ loads folded in)
  const double* fd;    // C v_e of the elements, or null
  const double* body;  // F_body, or null
  const double* m;
  const double* am;    // lumped alphaM mass, or null
  const double* dash;  // dampers outside the elements, or null
  const double* u;     // U(n)
  double* um;          // in: U(n-1); out: U(n+1)
  double* v;
  double* a;
  const int8_t* dir;   // the DOF direction of each equation
  int32_t gdir;        // the excitation's direction, or -1
  double ag;           // its acceleration at this step's load time
};

// The effective ground-motion load on equation q (ground_load, explicit_analysis.cpp:48-58):
// -m(q) a_g when q acts in the excitation's direction.
__device__ inline double ground_load(const CdArgs& g, int32_t q, double mq) {
  double r = 0.0;
  if (g.gdir >= 0 && g.dir[q] == g.gdir) r -= mq * g.ag;
  return r;
}

// RECORD: also the recorded velocity and acceleration. They feed no later displacement (the
// dampers act at the garbage velocity), so the loop writes them only in its last two steps: the
// last step's acceleration needs the velocity of the step before.
template <bool RECORD>
__global__ __launch_bounds__(kThreads) void cd_update_kernel(CdArgs g) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.n) return;
  const double uq = g.u[q], umq = g.um[q], mq = g.m[q];
  double fq = g.f[q];
  if (g.fd) fq += g.fd[q];
  if (g.body) fq -= g.body[q];
  double cq = 0.0;
  if (g.am) cq += g.am[q];
  if (g.dash) {
    const double dq = g.dash[q];
    fq += dq * (umq * -g.c2);  // the damper's force at the garbage velocity
    cq += dq;
  }
  const double ag = uq * (-2.0 * g.c3) + umq * g.c3;
  const double R = -fq - mq * ag + ground_load(g, q, mq);
  const double A = mq * g.c3 + cq * g.c2;
  const double U = R / A;
  g.um[q] = U;
  if constexpr (RECORD) {
    double Udot = U * 3.0;
    Udot += uq * -4.0;
    Udot += umq;
    Udot *= g.c2;
    const double Udotdot = (Udot - g.v[q]) / g.dt;
    g.v[q] = Udot;
    g.a[q] = Udotdot;
  }
}

// ExplicitDifference newStep (ed_predict, explicit_analysis.cpp:60-68): vh += dt a; u += dt vh.
__global__ __launch_bounds__(kThreads) void ed_predict_kernel(int32_t n, double dt, double* __restrict__ u,
                                                              double* __restrict__ vh, const double* __restrict__ a) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  vh[q] += a[q] * dt;
  u[q] += vh[q] * dt;
}

// ExplicitDifference solve and update (ed_correct, :70-85): a(n+1) = R / m with R = -f + the
// ground load; the recorded velocity vh + (dt/8)(3 a(n+1) + a(n)). The force f carries the
// elements' damping at vh; dampers outside the elements add dash * vh to it (no tangent term:
// the scheme is explicit in the velocity).
__global__ __launch_bounds__(kThreads) void ed_correct_kernel(CdArgs g) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= g.n) return;
  const double mq = g.m[q];
  double fq = g.f[q];
  if (g.fd) fq += g.fd[q];
  if (g.body) fq -= g.body[q];
  if (g.dash) fq += g.dash[q] * g.um[q];
  const double R = -fq + ground_load(g, q, mq);
  const double a_new = R / mq;
  const double t = 3.0 * a_new + g.a[q];
  g.v[q] = g.um[q] + t * (g.dt * 0.125);
  g.a[q] = a_new;
}

bool elements_damp(const fx_batch* b) {
  for (const DevParam& p : b->host_params)
    if (p.alphaM != 0.0 || p.betaK != 0.0 || p.betaKinit != 0.0 || p.betaKcomm != 0.0) return true;
  return false;
}

}  // namespace

int force64(fx_batch* b, const double* u, const double* v, double* f, double* scratch) {
  if (b->pimy) {  // the return map (and the damping force at v)
    if (int rc = pimy_force64(b, u, v, f, false)) return rc;
  } else {
    const int32_t n = b->neq;
    if (int rc = elements_apply64(b, fx_coeffs{1.0, 0.0, 0.0, 0}, u, f)) return rc;
    const double* cv = nullptr;
    if (v && elements_damp(b)) {
      if (int rc = elements_apply64(b, fx_coeffs{0.0, 1.0, 0.0, 0}, v, scratch)) return rc;
      cv = scratch;
    }
    if (cv || b->body)
      hipLaunchKernelGGL(force_combine_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, b->ctx->stream, n, f, cv,
                         b->body, f);
    FX_HIP(b->ctx, hipGetLastError());
  }
  // the nodal masses' Rayleigh damping force alphaM m v (internal_force.cpp:366-374)
  if (v && b->nodal_am) return nodal_damping_force64(b, v, f);
  return FX_OK;
}

int cd_run(fx_batch* b, const fx_cd_desc& d, double* u, double* um, double* v, double* a) {
  fx_ctx* ctx = b->ctx;
  const int32_t n = b->neq;
  if (n == 0 || d.nsteps <= 0) return FX_OK;
  double* W = b->work;
  const size_t N = static_cast<size_t>(n);
  // state in the solver numbering
  double *U = u, *Um = um, *V = v, *Acc = a, *dash = const_cast<double*>(d.dash);
  if (b->internal) {
    U = W;
    Um = W + N;
    V = W + 2 * N;
    Acc = W + 3 * N;
    if (int rc = to_internal<double>(b, u, U)) return rc;
    if (int rc = to_internal<double>(b, um, Um)) return rc;
    if (int rc = to_internal<double>(b, v, V)) return rc;
    if (int rc = to_internal<double>(b, a, Acc)) return rc;
    if (d.dash) {
      dash = W + 6 * N;
      if (int rc = to_internal<double>(b, d.dash, dash)) return rc;
    }
  }
  double* f = W + 4 * N;
  double* ve = W + 5 * N;
  double* fd = b->internal ? W + 7 * N : W + 6 * N;
  // the elements' damping force at the garbage velocity: by the operator (C v_e) for elastic
  // batches, inside the return-map force for PIMY batches
  // (the nodal masses' alphaM term joins the operator's, or follows the return-map force)
  const bool pdamp = b->pimy && elements_damp(b);
  const bool edamp = elements_damp(b) || b->nodal_am, damp = !b->pimy && edamp;
  const bool ed = d.scheme == 1;  // ExplicitDifference: um holds the element velocity vh
  const double c2 = 0.5 / d.dt, c3 = 1.0 / (d.dt * d.dt);
  // the excitation, evaluated where the reference's stage loop evaluates its loads: at t(n) for
  // CentralDifference, at t(n+1) for ExplicitDifference (explicit_analysis.cpp:699-722)
  int32_t loc = d.ground_series && d.ground_series->loc ? *d.ground_series->loc : 0;
  double time = d.time;
  auto ground_at = [&](int32_t k, double t) {
    if (d.ground_dir < 0) return 0.0;
    if (d.ground_series) return series_value(*d.ground_series, d.ground_factor, t, loc);
    return d.ground_accel[k];
  };
  PointLoads pl{};
  pl.n = d.nload;
  for (int j = 0; j < d.nload; ++j) pl.eq[j] = b->internal ? b->internal_of[d.load_eq[j]] : d.load_eq[j];
  CdArgs g{};
  g.n = n;
  g.c2 = c2;
  g.c3 = c3;
  g.dt = d.dt;
  g.f = f;
  g.fd = damp ? fd : nullptr;
  g.body = b->pimy ? nullptr : b->body;  // the return-map force carries the body force
  g.m = b->lumped_m;
  g.am = b->lumped_am;
  g.dash = dash;
  g.dir = b->eq_dir_dev;
  g.gdir = d.ground_dir;
  // node recorders after every step (explicit_analysis.cpp:724-731): u and the recorded v, a
  const bool rec = recorder_active(d.recorder), rec_rates = recorder_reads_rates(d.recorder);
  const int32_t* rmap = nullptr;
  if (rec)
    if (int rc = batch_internal_map(b, &rmap)) return rc;
  if (ed) {
    double* VH = Um;
    for (int32_t k = 0; k < d.nsteps; ++k) {
      hipLaunchKernelGGL(ed_predict_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, ctx->stream, n, d.dt, U, VH, Acc);
      time += d.dt;
      g.ag = ground_at(k, time);
      if (damp)
        if (int rc = operator_apply64(b, fx_coeffs{0.0, 1.0, 0.0, 0}, VH, fd)) return rc;
      if (b->pimy) {
        if (int rc = pimy_force64(b, U, pdamp ? VH : nullptr, f, false)) return rc;
        if (b->nodal_am)
          if (int rc = nodal_damping_force64(b, VH, f)) return rc;
      } else if (int rc = operator_apply64(b, fx_coeffs{1.0, 0.0, 0.0, 0}, U, f)) {
        return rc;
      }
      if (pl.n > 0) {
        for (int j = 0; j < pl.n; ++j) pl.val[j] = d.load_hist[static_cast<size_t>(k) * pl.n + j];
        hipLaunchKernelGGL(point_load_kernel, dim3(1), dim3(64), 0, ctx->stream, pl, f);
      }
      g.u = U;
      g.um = VH;
      g.v = V;
      g.a = Acc;
      hipLaunchKernelGGL(ed_correct_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, ctx->stream, g);
      if (b->pimy) std::swap(b->state_c, b->state_t);  // commit_state
      if (rec)
        if (int rc = recorder_record(d.recorder, time, n, U, V, Acc, rmap)) return rc;
    }
  }
  for (int32_t k = 0; !ed && k < d.nsteps; ++k) {
    g.ag = ground_at(k, time);
    time += d.dt;
    if (edamp)
      hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, ctx->stream, n, -c2, Um, ve);
    if (damp)
      if (int rc = operator_apply64(b, fx_coeffs{0.0, 1.0, 0.0, 0}, ve, fd)) return rc;
    if (b->pimy) {
      if (int rc = pimy_force64(b, U, pdamp ? ve : nullptr, f, false)) return rc;
      if (b->nodal_am)
        if (int rc = nodal_damping_force64(b, ve, f)) return rc;
    } else if (int rc = operator_apply64(b, fx_coeffs{1.0, 0.0, 0.0, 0}, U, f)) {
      return rc;
    }
    if (pl.n > 0) {
      for (int j = 0; j < pl.n; ++j) pl.val[j] = d.load_hist[static_cast<size_t>(k) * pl.n + j];
      hipLaunchKernelGGL(point_load_kernel, dim3(1), dim3(64), 0, ctx->stream, pl, f);
    }
    g.u = U;
    g.um = Um;
    g.v = V;
    g.a = Acc;
    // a record of v or a needs this step's (and so the step before's v): the full update every
    // step, as the reference's cd_solve always does
    if (rec_rates || k >= d.nsteps - 2)
      hipLaunchKernelGGL(cd_update_kernel<true>, dim3(blocks_for(n)), dim3(kThreads), 0, ctx->stream, g);
    else
      hipLaunchKernelGGL(cd_update_kernel<false>, dim3(blocks_for(n)), dim3(kThreads), 0, ctx->stream, g);
    std::swap(U, Um);  // U(n+1) sits in the old U(n-1) buffer
    if (b->pimy) std::swap(b->state_c, b->state_t);  // commit_state: the trial history becomes committed
    if (rec)
      if (int rc = recorder_record(d.recorder, time, n, U, V, Acc, rmap)) return rc;
  }
  FX_HIP(ctx, hipGetLastError());
  // back to the caller's buffers: u = U(n+1), um = U(n)
  if (b->internal) {
    if (int rc = to_external<double>(b, U, u)) return rc;
    if (int rc = to_external<double>(b, Um, um)) return rc;
    if (int rc = to_external<double>(b, V, v)) return rc;
    if (int rc = to_external<double>(b, Acc, a)) return rc;
  } else if (!ed && U != u) {  // an odd number of steps: the roles are swapped
    if (int rc = copy64(ctx, n, U, ve)) return rc;
    if (int rc = copy64(ctx, n, Um, U)) return rc;
    if (int rc = copy64(ctx, n, ve, Um)) return rc;
  }
  if (d.ground_series && d.ground_series->loc) *d.ground_series->loc = loc;
  FX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FX_OK;
}

}  // namespace fx
