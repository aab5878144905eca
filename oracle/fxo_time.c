/*
 * fxo internal force, stable time step and the CentralDifference loop — TEST INFRASTRUCTURE
 * (the oracle; see fxo.h).
 *
 * Restates:
 *   InternalForce::compute, flat force kernel, slot emit + gather
 *                                  src/exec/internal_force.cpp:354-377,
 *                                  src/exec/element_kernels.hpp:198-235, assembly.cpp:64-95
 *   StdBrick::max_omega_sq_with    src/physics/elements/std_brick.hpp:658-687
 *   max_eigenvalue_symmetric       src/physics/common/sym_eigen.hpp:11-49 (cyclic Jacobi,
 *                                  Golub & Van Loan, Matrix Computations, Algorithm 8.4.1)
 *   stable_dt                      src/physics/explicit_stability.hpp:15-32
 *   CentralDifference stage loop   src/exec/explicit_analysis.cpp:87-117, 661-662, 712-722
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fxo.h"

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

void fxo_force(const fxo_compiled* c, const fxo_elastic* params, const double* u, const double* v, double* f,
               int nthreads) {
  set_threads(nthreads);
  const int32_t N = c->nnode, P = c->padded, E = c->nelem;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    double xe[8][3], ue[24], ve[24], fe[24];
    for (int a = 0; a < 8; ++a) {
      const int32_t n = c->node[(int64_t)a * P + e];
      for (int d = 0; d < 3; ++d) xe[a][d] = c->coords[(int64_t)d * N + n];
    }
    for (int k = 0; k < 24; ++k) {
      const int32_t q = c->eq[(int64_t)k * P + e];
      ue[k] = q >= 0 ? u[q] : 0.0;
      ve[k] = (q >= 0 && v) ? v[q] : 0.0;
    }
    fxo_brick_force_elastic(&params[c->param[e]], (const double(*)[3])xe, ue, ve, fe, NULL);
    for (int k = 0; k < 24; ++k) c->slots[(int64_t)k * P + e] = fe[k];
  }
  const int32_t neq = c->neq;
#pragma omp parallel for schedule(static)
  for (int32_t q = 0; q < neq; ++q) {
    double sum = 0.0;
    for (int64_t j = c->gather_start[q]; j < c->gather_start[q + 1]; ++j) sum += c->slots[c->gather_list[j]];
    f[q] = sum;
  }
}

double fxo_max_eigenvalue_symmetric24(double A[24][24]) {
  const int n = 24;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < n; ++i) {
      diag += A[i][i] * A[i][i];
      for (int j = i + 1; j < n; ++j) off += A[i][j] * A[i][j];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p][q];
        if (apq == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0);
        const double sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = cs * akp - sn * akq;
          A[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = cs * apk - sn * aqk;
          A[q][k] = sn * apk + cs * aqk;
        }
        A[p][q] = A[q][p] = 0.0;
      }
  }
  double lmax = A[0][0];
  for (int i = 1; i < n; ++i)
    if (A[i][i] > lmax) lmax = A[i][i];
  return lmax;
}

double fxo_brick_max_omega_sq(const fxo_elastic* p, const double xe[8][3]) {
  if (p->element == 1) return fxo_ssp_max_omega_sq(p, xe);
  double m[8], K[576], fe[24], zero[24];
  memset(zero, 0, sizeof(zero));
  fxo_brick_lumped_mass(p->rho, xe, m);
  fxo_elastic q = *p;
  q.b[0] = q.b[1] = q.b[2] = 0.0;
  fxo_brick_force_elastic(&q, xe, zero, zero, fe, K);
  double A[24][24], s[8];
  for (int j = 0; j < 8; ++j) {
    if (!(m[j] > 0.0)) return 0.0; /* massless: no limit */
    s[j] = 1.0 / sqrt(m[j]);
  }
  for (int a = 0; a < 24; ++a)
    for (int c = 0; c < 24; ++c) A[a][c] = K[a * 24 + c] * s[a / 3] * s[c / 3];
  return fxo_max_eigenvalue_symmetric24(A);
}

double fxo_stable_dt(int scheme, const fxo_rayleigh* damp, double omega_sq) {
  if (!(omega_sq > 0.0)) return INFINITY;
  const double omega = sqrt(omega_sq);
  if (scheme == 0) return 2.0 / omega; /* CentralDifference */
  const double beta = damp->betaK + damp->betaKinit + damp->betaKcomm;
  const double cc = damp->alphaM + beta * omega_sq;
  return 4.0 / (sqrt(cc * cc + 4.0 * omega_sq) + cc);
}

double fxo_critical_dt(const fxo_compiled* c, const fxo_elastic* params, int scheme) {
  const int32_t N = c->nnode, P = c->padded, E = c->nelem;
  double dt = INFINITY;
  for (int32_t e = 0; e < E; ++e) {
    double xe[8][3];
    for (int a = 0; a < 8; ++a) {
      const int32_t n = c->node[(int64_t)a * P + e];
      for (int d = 0; d < 3; ++d) xe[a][d] = c->coords[(int64_t)d * N + n];
    }
    const fxo_elastic* p = &params[c->param[e]];
    const double s = fxo_stable_dt(scheme, &p->damp, fxo_brick_max_omega_sq(p, (const double(*)[3])xe));
    if (s < dt) dt = s;
  }
  return dt;
}

void fxo_explicit_run(const fxo_compiled* c, const fxo_elastic* params, const double* m, const double* cm,
                      const double* dash, int32_t nload, const int32_t* load_eq, const double* load_hist, double dt,
                      int32_t nsteps, int scheme, int ground_dir, const double* ground_accel, double* u, double* um,
                      double* v, double* a, int nthreads) {
  const int32_t n = c->neq;
  const double c2 = 0.5 / dt, c3 = 1.0 / (dt * dt);
  double* f = (double*)malloc(((size_t)n + 1) * sizeof(double));
  double* ve = (double*)malloc(((size_t)n + 1) * sizeof(double));
  double* P = (double*)calloc((size_t)n + 1, sizeof(double));
  signed char* dir = (signed char*)malloc((size_t)n + 1);
  for (int32_t q = 0; q < n; ++q) dir[q] = -1;
  for (int32_t i = 0; i < 3 * c->nnode; ++i)
    if (c->node_eq[i] >= 0) dir[c->node_eq[i]] = (signed char)(i % 3);
  for (int32_t k = 0; k < nsteps; ++k) {
    for (int32_t j = 0; j < nload; ++j) P[load_eq[j]] = load_hist[(int64_t)k * nload + j];
    const double ag_k = ground_dir >= 0 ? ground_accel[k] : 0.0;
    if (scheme == 1) {
      /* ExplicitDifference (explicit_analysis.cpp:60-85): vh (in um) += dt a; u += dt vh; the
       * force at vh; a = R / m; the recorded velocity vh + dt/8 (3 a(n+1) + a(n)) */
      for (int32_t q = 0; q < n; ++q) {
        um[q] += a[q] * dt;
        u[q] += um[q] * dt;
      }
      fxo_force(c, params, u, um, f, nthreads);
      if (dash)
        for (int32_t q = 0; q < n; ++q) f[q] += dash[q] * um[q];
      for (int32_t q = 0; q < n; ++q) {
        double g = 0.0; /* ground_load (:48-58) */
        if (dir[q] == ground_dir) g -= m[q] * ag_k;
        const double R = -f[q] + P[q] + g;
        const double a_new = R / m[q];
        const double t = 3.0 * a_new + a[q];
        v[q] = um[q] + t * (dt * 0.125);
        a[q] = a_new;
      }
      continue;
    }
    /* CentralDifference: loads at t(n); the garbage velocity for the element damping force */
    for (int32_t q = 0; q < n; ++q) ve[q] = um[q] * -c2;
    fxo_force(c, params, u, ve, f, nthreads);
    if (dash)
      for (int32_t q = 0; q < n; ++q) f[q] += dash[q] * ve[q];
    for (int32_t q = 0; q < n; ++q) {
      double g = 0.0;
      if (dir[q] == ground_dir) g -= m[q] * ag_k;
      const double ag = u[q] * (-2.0 * c3) + um[q] * c3;
      const double R = -f[q] - m[q] * ag + P[q] + g;
      const double A = m[q] * c3 + (cm ? cm[q] : 0.0) * c2;
      const double U = R / A;
      double Udot = U * 3.0;
      Udot += u[q] * -4.0;
      Udot += um[q];
      Udot *= c2;
      const double Udotdot = (Udot - v[q]) / dt;
      um[q] = u[q];
      u[q] = U;
      v[q] = Udot;
      a[q] = Udotdot;
    }
  }
  free(f);
  free(ve);
  free(P);
  free(dir);
}

void fxo_cd_run(const fxo_compiled* c, const fxo_elastic* params, const double* m, const double* cm,
                const double* dash, int32_t nload, const int32_t* load_eq, const double* load_hist, double dt, int32_t nsteps,
                double* u, double* um, double* v, double* a, int nthreads) {
  fxo_explicit_run(c, params, m, cm, dash, nload, load_eq, load_hist, dt, nsteps, 0, -1, NULL, u, um, v, a, nthreads);
}
