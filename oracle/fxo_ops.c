/*
 * fxo global operator, reductions and CG — TEST INFRASTRUCTURE (the oracle; see fxo.h).
 *
 * Restates:
 *   element gather and zero rules        src/exec/operator_kernels.hpp:75-141
 *   slot emit + deterministic gather     src/exec/assembly.cpp:64-95 with the pair order of
 *                                        src/exec/internal_force.cpp:121-134
 *   InternalForce::apply_operator / operator_diagonal   internal_force.cpp:545-599
 *   lumped mass                          src/compile/batches.cpp:204-224
 *   VectorOps::dot (chunk 2048, lanes)   src/solvers/vector_ops.cpp:19-73, host_dot :105-124
 *   ConjugateGradient::solve             src/solvers/cg.cpp:24-72
 *   Jacobi                               src/solvers/jacobi.cpp:10-26
 * The OpenMP parallel loops write disjoint outputs and keep every sum's order, so the result
 * is the same bits for any thread count (as the reference's test_vector_ops/test_internal_force
 * require).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fxo.h"

static void gather_xe(const fxo_compiled* c, int32_t e, double xe[8][3]) {
  const int32_t N = c->nnode, P = c->padded;
  for (int a = 0; a < 8; ++a) {
    const int32_t n = c->node[(int64_t)a * P + e];
    for (int d = 0; d < 3; ++d) xe[a][d] = c->coords[(int64_t)d * N + n];
  }
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* y(q) = sum of q's slots in (element, DOF) order: the SlotGather result */
static void gather(const fxo_compiled* c, double* y) {
  const int32_t neq = c->neq;
#pragma omp parallel for schedule(static)
  for (int32_t q = 0; q < neq; ++q) {
    double sum = 0.0;
    for (int64_t j = c->gather_start[q]; j < c->gather_start[q + 1]; ++j) sum += c->slots[c->gather_list[j]];
    y[q] = sum;
  }
}

/* InternalForce::mass_times_ground (internal_force.cpp:601-626; the unit_dir source of
 * operator_kernels.hpp:75-141): y = M r with r the rigid unit translation in direction dir at every
 * element DOF, fixed ones included; the element masses of elastic or PIMY parameter sets. */
void fxo_mass_times_ground(const fxo_compiled* c, const fxo_elastic* elastic, const fxo_pimy_elem* pimy, int dir,
                           double* y, int nthreads) {
  set_threads(nthreads);
  const int32_t P = c->padded, E = c->nelem;
  const fxo_coeffs mass = {0.0, 0.0, 1.0, 0};
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    double xe[8][3], ve[24], ye[24], D[36];
    gather_xe(c, e, xe);
    for (int a = 0; a < 24; ++a) ve[a] = a % 3 == dir ? 1.0 : 0.0;
    memset(D, 0, sizeof D);
    if (pimy) {
      double Dgp[8 * 36];
      memset(Dgp, 0, sizeof Dgp);
      fxo_brick_operator_apply_gp(&pimy[c->param[e]], &mass, xe, Dgp, NULL, ve, ye);
    } else {
      fxo_brick_operator_apply(&elastic[c->param[e]], &mass, xe, D, ve, ye);
    }
    for (int a = 0; a < 24; ++a) c->slots[(int64_t)a * P + e] = ye[a];
  }
  gather(c, y);
}

void fxo_apply(const fxo_compiled* c, const fxo_elastic* params, const fxo_coeffs* k,
               const double* x, double* y, int nthreads) {
  set_threads(nthreads);
  const int32_t P = c->padded, E = c->nelem;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    const fxo_elastic* p = &params[c->param[e]];
    double xe[8][3], ve[24], ye[24], D[6][6];
    gather_xe(c, e, xe);
    for (int a = 0; a < 24; ++a) {
      const int32_t q = c->eq[(int64_t)a * P + e];
      ve[a] = q >= 0 ? x[q] : 0.0;
    }
    fxo_elastic_tangent(p->E, p->nu, D);
    fxo_brick_operator_apply(p, k, xe, &D[0][0], ve, ye);
    for (int a = 0; a < 24; ++a) c->slots[(int64_t)a * P + e] = ye[a];
  }
  gather(c, y);
}

void fxo_diagonal(const fxo_compiled* c, const fxo_elastic* params, const fxo_coeffs* k, double* d,
                  int nthreads) {
  set_threads(nthreads);
  const int32_t P = c->padded, E = c->nelem;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    const fxo_elastic* p = &params[c->param[e]];
    double xe[8][3], de[24], D[6][6];
    int32_t eqe[24];
    gather_xe(c, e, xe);
    for (int a = 0; a < 24; ++a) eqe[a] = c->eq[(int64_t)a * P + e];
    fxo_elastic_tangent(p->E, p->nu, D);
    fxo_brick_operator_diagonal(p, k, xe, &D[0][0], eqe, de);
    for (int a = 0; a < 24; ++a) c->slots[(int64_t)a * P + e] = de[a];
  }
  gather(c, d);
}

void fxo_lumped_mass(const fxo_compiled* c, const fxo_elastic* params, double* m, double* am) {
  const int32_t P = c->padded;
  for (int32_t q = 0; q < c->neq; ++q) {
    m[q] = 0.0;
    if (am) am[q] = 0.0;
  }
  for (int32_t e = 0; e < c->nelem; ++e) {
    const fxo_elastic* p = &params[c->param[e]];
    double xe[8][3], mj[8];
    gather_xe(c, e, xe);
    fxo_brick_lumped_mass(p->rho, xe, mj);
    for (int k = 0; k < 24; ++k) {
      const int32_t q = c->eq[(int64_t)k * P + e];
      if (q < 0) continue;
      m[q] += mj[k / 3];
      if (am) am[q] += p->damp.alphaM * mj[k / 3];
    }
  }
}

void fxo_body_load(const fxo_compiled* c, const fxo_elastic* params, double* b) {
  /* b = -f(u = 0, v = 0): the element force with zero fields is -body (resid -= body) */
  const int32_t P = c->padded, E = c->nelem;
  double zero[24];
  for (int a = 0; a < 24; ++a) zero[a] = 0.0;
  for (int32_t e = 0; e < E; ++e) {
    fxo_elastic p = params[c->param[e]];
    p.damp.alphaM = p.damp.betaK = p.damp.betaKinit = p.damp.betaKcomm = 0.0;
    double xe[8][3], fe[24];
    gather_xe(c, e, xe);
    fxo_brick_force_elastic(&p, xe, zero, zero, fe, NULL);
    for (int a = 0; a < 24; ++a) c->slots[(int64_t)a * P + e] = -fe[a];
  }
  gather(c, b);
}

/* ---- VectorOps -------------------------------------------------------------------------------- */
#define CHUNK 2048
static void reduce_level(int64_t n, const double* a, const double* b, int lanes, double* out) {
  const int64_t nch = (n + CHUNK - 1) / CHUNK;
  double* lane = (double*)malloc((size_t)lanes * sizeof(double));
  for (int64_t ch = 0; ch < nch; ++ch) {
    const int64_t begin = ch * CHUNK, len = (n - begin < CHUNK) ? n - begin : CHUNK;
    for (int l = 0; l < lanes; ++l) {
      double s = 0.0;
      for (int64_t k = l; k < len; k += lanes) s += b ? a[begin + k] * b[begin + k] : a[begin + k];
      lane[l] = s;
    }
    double total = 0.0;
    for (int l = 0; l < lanes; ++l) total += lane[l];
    out[ch] = total;
  }
  free(lane);
}

double fxo_dot(int64_t n, const double* x, const double* y, int lanes) {
  if (n == 0) return 0.0;
  if (lanes < 1) lanes = 1;
  int64_t m = (n + CHUNK - 1) / CHUNK;
  double* p = (double*)malloc((size_t)m * sizeof(double));
  double* q = (double*)malloc((size_t)m * sizeof(double));
  reduce_level(n, x, y, lanes, p);
  while (m > 1) {
    reduce_level(m, p, NULL, lanes, q);
    m = (m + CHUNK - 1) / CHUNK;
    double* t = p;
    p = q;
    q = t;
  }
  const double r = p[0];
  free(p);
  free(q);
  return r;
}

/* ---- CG + Jacobi ----------------------------------------------------------------------------- */
int fxo_cg(const fxo_compiled* c, const fxo_elastic* params, const fxo_coeffs* k, const double* b,
           double* x, double rtol, double atol, int32_t max_it, int lanes, int nthreads,
           fxo_solve_result* res) {
  const int64_t n = c->neq;
  memset(res, 0, sizeof(*res));
  double* r = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* z = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* p = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* ap = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* inv = (double*)malloc((size_t)(n + 1) * sizeof(double));
  int status = 0;
  /* Jacobi::setup */
  fxo_diagonal(c, params, k, inv, nthreads);
  for (int64_t i = 0; i < n; ++i) {
    if (inv[i] == 0.0) status = -1;
    inv[i] = 1.0 / inv[i];
  }
  if (status == 0) {
    res->rhs_norm = sqrt(fxo_dot(n, b, b, lanes));
    fxo_apply(c, params, k, x, ap, nthreads);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i];
    for (int64_t i = 0; i < n; ++i) r[i] += -1.0 * ap[i];
    double rnorm = sqrt(fxo_dot(n, r, r, lanes));
    res->initial_residual = res->final_residual = rnorm;
    const double tol = fmax(rtol * res->rhs_norm, atol);
    if (rnorm <= tol) {
      res->converged = 1;
    } else {
      for (int64_t i = 0; i < n; ++i) z[i] = inv[i] * r[i];
      for (int64_t i = 0; i < n; ++i) p[i] = z[i];
      double rz = fxo_dot(n, r, z, lanes);
      for (int32_t it = 1; it <= max_it; ++it) {
        fxo_apply(c, params, k, p, ap, nthreads);
        const double pap = fxo_dot(n, p, ap, lanes);
        if (!(pap > 0.0)) {
          res->indefinite = 1;
          res->iterations = it - 1;
          break;
        }
        const double alpha = rz / pap;
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] += -alpha * ap[i];
        rnorm = sqrt(fxo_dot(n, r, r, lanes));
        res->iterations = it;
        res->final_residual = rnorm;
        if (rnorm <= tol) {
          res->converged = 1;
          break;
        }
        for (int64_t i = 0; i < n; ++i) z[i] = inv[i] * r[i];
        const double rz_new = fxo_dot(n, r, z, lanes);
        const double beta = rz_new / rz;
        rz = rz_new;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      }
    }
  }
  free(r);
  free(z);
  free(p);
  free(ap);
  free(inv);
  return status;
}

/* ---- PressureIndependMultiYield batches (points e * 8 + gp, batch order; history point-major) */
void fxo_pimy_force(const fxo_compiled* c, const fxo_pimy_elem* params, const double* C, double* T,
                    const double* u, const double* v, double* f, double* D, int nthreads) {
  set_threads(nthreads);
  const int32_t P = c->padded, E = c->nelem;
  const int ns = fxo_pimy_nstate();
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    double xe[8][3], ue[24], ve[24], fe[24];
    gather_xe(c, e, xe);
    for (int k = 0; k < 24; ++k) {
      const int32_t q = c->eq[(int64_t)k * P + e];
      ue[k] = q >= 0 ? u[q] : 0.0;
      ve[k] = (v && q >= 0) ? v[q] : 0.0;
    }
    const double* Cp[8];
    double* Tp[8];
    for (int gp = 0; gp < 8; ++gp) {
      Cp[gp] = C + ((int64_t)e * 8 + gp) * ns;
      Tp[gp] = T + ((int64_t)e * 8 + gp) * ns;
    }
    fxo_brick_force_pimy(&params[c->param[e]], (const double(*)[3])xe, ue, v ? ve : NULL, Cp, Tp, fe,
                         D ? D + (int64_t)e * 8 * 36 : NULL);
    for (int k = 0; k < 24; ++k) c->slots[(int64_t)k * P + e] = fe[k];
  }
  gather(c, f);
}

/* the committed history pointers of element e's points, or NULLs */
static void point_ptrs(const double* C, int32_t e, const double** Cp) {
  const int ns = fxo_pimy_nstate();
  for (int gp = 0; gp < 8; ++gp) Cp[gp] = C ? C + ((int64_t)e * 8 + gp) * ns : NULL;
}

void fxo_apply_gp(const fxo_compiled* c, const fxo_pimy_elem* params, const fxo_coeffs* k, const double* D,
                  const double* C, const double* x, double* y, int nthreads) {
  set_threads(nthreads);
  const int32_t P = c->padded, E = c->nelem;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    double xe[8][3], ve[24], ye[24];
    gather_xe(c, e, xe);
    for (int a = 0; a < 24; ++a) {
      const int32_t q = c->eq[(int64_t)a * P + e];
      ve[a] = q >= 0 ? x[q] : 0.0;
    }
    const double* Cp[8];
    point_ptrs(C, e, Cp);
    fxo_brick_operator_apply_gp(&params[c->param[e]], k, (const double(*)[3])xe, D + (int64_t)e * 8 * 36, Cp, ve, ye);
    for (int a = 0; a < 24; ++a) c->slots[(int64_t)a * P + e] = ye[a];
  }
  gather(c, y);
}

void fxo_diagonal_gp(const fxo_compiled* c, const fxo_pimy_elem* params, const fxo_coeffs* k, const double* D,
                     const double* C, double* d, int nthreads) {
  set_threads(nthreads);
  const int32_t P = c->padded, E = c->nelem;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    double xe[8][3], de[24];
    int32_t eq[24];
    gather_xe(c, e, xe);
    for (int a = 0; a < 24; ++a) eq[a] = c->eq[(int64_t)a * P + e];
    const double* Cp[8];
    point_ptrs(C, e, Cp);
    fxo_brick_operator_diagonal_gp(&params[c->param[e]], k, (const double(*)[3])xe, D + (int64_t)e * 8 * 36, Cp, eq,
                                   de);
    for (int a = 0; a < 24; ++a) c->slots[(int64_t)a * P + e] = de[a];
  }
  gather(c, d);
}
