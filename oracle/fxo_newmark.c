/*
 * fxo implicit Newmark stage with Newton iterations — TEST INFRASTRUCTURE (the oracle; see fxo.h).
 *
 * Restates src/exec/implicit_stage.cpp: newmark_predict :29-40, newmark_update :43-52,
 * form_residual :55-68, the integrator constants and operator coefficients :150-159, 283-286,
 * update_tangent :180-193, form_unbalance :212-219, solve :223-238 (Newton accepts an increment
 * whose solve reduced the residual by 1e-6), test_value :239-246, the Newton attempt :274-372,
 * recover :385-402 (retries, substeps) and the stage loop :404-418, static LoadControl stages, the
 * Linear and ModifiedNewton (-initial, -factoronce) algorithms :118-372, the solver KrylovSolver
 * picks from the operator's symmetry (krylov.cpp:10-42, matrix_free_operator.cpp:5-16: CG with
 * the GMRES fallback for an SPD operator, GMRES otherwise), the ground acceleration at
 * each attempt's end time (GroundMotions::evaluate, ground_motion.cpp:17-33, 121-145), commit
 * :365-372; the PCG of src/solvers/cg.cpp:24-72 with
 * Jacobi (jacobi.cpp:10-26) on a generic system (elastic operator, or the stored tangents of
 * PressureIndependMultiYield batches).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fxo.h"

static void sys_force(fxo_system* s, const double* u, const double* v, double* f) {
  if (s->pimy) fxo_pimy_force(s->c, s->pimy, s->C, s->T, u, v, f, NULL, s->nthreads);
  else fxo_force(s->c, s->elastic, u, v, f, s->nthreads);
}
static void sys_store(fxo_system* s, const double* u, double* scratch) {
  if (s->pimy) fxo_pimy_force(s->c, s->pimy, s->C, s->T, u, NULL, scratch, s->D, s->nthreads);
}
static void sys_apply(fxo_system* s, const fxo_coeffs* k, const double* x, double* y) {
  if (s->pimy) fxo_apply_gp(s->c, s->pimy, k, s->D, s->C, x, y, s->nthreads);
  else fxo_apply(s->c, s->elastic, k, x, y, s->nthreads);
}
static void sys_diagonal(fxo_system* s, const fxo_coeffs* k, double* d) {
  if (s->pimy) fxo_diagonal_gp(s->c, s->pimy, k, s->D, s->C, d, s->nthreads);
  else fxo_diagonal(s->c, s->elastic, k, d, s->nthreads);
}

int fxo_cg_system(fxo_system* s, const fxo_coeffs* k, const double* b, double* x, double rtol, double atol,
                  int32_t max_it, int lanes, fxo_solve_result* res) {
  const int64_t n = s->c->neq;
  memset(res, 0, sizeof(*res));
  double* r = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* z = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* p = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* ap = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* inv = (double*)malloc((size_t)(n + 1) * sizeof(double));
  int status = 0;
  sys_diagonal(s, k, inv);
  for (int64_t i = 0; i < n; ++i) {
    if (inv[i] == 0.0) status = -1;
    inv[i] = 1.0 / inv[i];
  }
  if (status == 0) {
    res->rhs_norm = sqrt(fxo_dot(n, b, b, lanes));
    sys_apply(s, k, x, ap);
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
        sys_apply(s, k, p, ap);
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

/* FGMRES(m) with CGS2 and right Jacobi preconditioning (gmres.cpp:32-123; BlockOps::dots is a
 * VectorOps dot per column, combine adds sum_c coef_c V(:, c) per row, block_ops.cpp:59-90) */
int fxo_gmres_system(fxo_system* s, const fxo_coeffs* k, const double* b, double* x, double rtol, double atol,
                     int32_t max_it, int32_t m, int lanes, fxo_solve_result* res) {
  const int64_t n = s->c->neq;
  memset(res, 0, sizeof(*res));
  double* V = (double*)malloc((size_t)(n * (m + 1) + 1) * sizeof(double));
  double* Z = (double*)malloc((size_t)(n * m + 1) * sizeof(double));
  double* w = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* r = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* inv = (double*)malloc((size_t)(n + 1) * sizeof(double));
  double* H = (double*)calloc((size_t)((m + 1) * m), sizeof(double));
  double *cs = (double*)calloc((size_t)m + 1, sizeof(double)), *sn = (double*)calloc((size_t)m + 1, sizeof(double));
  double *g = (double*)calloc((size_t)m + 2, sizeof(double)), *h = (double*)calloc((size_t)m + 2, sizeof(double));
  double *h2 = (double*)calloc((size_t)m + 2, sizeof(double)), *y = (double*)calloc((size_t)m + 2, sizeof(double));
#define HIJ(i, j) H[(size_t)(j) * (m + 1) + (i)]
  int status = 0;
  sys_diagonal(s, k, inv);
  for (int64_t i = 0; i < n; ++i) {
    if (inv[i] == 0.0) status = -1;
    inv[i] = 1.0 / inv[i];
  }
  res->rhs_norm = sqrt(fxo_dot(n, b, b, lanes));
  const double tol = fmax(rtol * res->rhs_norm, atol);
  int first = 1;
  while (status == 0) {
    sys_apply(s, k, x, w);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i];
    for (int64_t i = 0; i < n; ++i) r[i] += -1.0 * w[i];
    const double beta = sqrt(fxo_dot(n, r, r, lanes));
    if (first) res->initial_residual = beta;
    first = 0;
    res->final_residual = beta;
    if (beta <= tol) {
      res->converged = 1;
      break;
    }
    if (res->iterations >= max_it) break;
    for (int64_t i = 0; i < n; ++i) V[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) V[i] += (1.0 / beta) * r[i];
    for (int i = 0; i <= m; ++i) g[i] = 0.0;
    g[0] = beta;
    int j = 0;
    for (; j < m && res->iterations < max_it; ++j) {
      double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n;
      for (int64_t i = 0; i < n; ++i) zj[i] = inv[i] * vj[i];
      sys_apply(s, k, zj, w);
      for (int pass = 0; pass < 2; ++pass) {
        double* hh = pass == 0 ? h : h2;
        for (int c = 0; c <= j; ++c) hh[c] = fxo_dot(n, V + (size_t)c * n, w, lanes);
        for (int64_t i = 0; i < n; ++i) {
          double acc = 0.0;
          for (int c = 0; c <= j; ++c) acc += (-1.0 * hh[c]) * V[(size_t)c * n + i];
          w[i] += acc;
        }
      }
      for (int i = 0; i <= j; ++i) HIJ(i, j) = h[i] + h2[i];
      const double hn = sqrt(fxo_dot(n, w, w, lanes));
      HIJ(j + 1, j) = hn;
      if (hn > 0.0) {
        double* vn = V + (size_t)(j + 1) * n;
        for (int64_t i = 0; i < n; ++i) vn[i] = 0.0;
        for (int64_t i = 0; i < n; ++i) vn[i] += (1.0 / hn) * w[i];
      }
      for (int i = 0; i < j; ++i) {
        const double a = HIJ(i, j), c = HIJ(i + 1, j);
        HIJ(i, j) = cs[i] * a + sn[i] * c;
        HIJ(i + 1, j) = -sn[i] * a + cs[i] * c;
      }
      const double a = HIJ(j, j), c = HIJ(j + 1, j);
      const double d = hypot(a, c);
      cs[j] = d > 0.0 ? a / d : 1.0;
      sn[j] = d > 0.0 ? c / d : 0.0;
      HIJ(j, j) = d;
      HIJ(j + 1, j) = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++res->iterations;
      res->final_residual = fabs(g[j + 1]);
      if (res->final_residual <= tol || hn == 0.0) {
        ++j;
        break;
      }
    }
    for (int i = j - 1; i >= 0; --i) {
      double sum = g[i];
      for (int kk = i + 1; kk < j; ++kk) sum -= HIJ(i, kk) * y[kk];
      y[i] = sum / HIJ(i, i);
    }
    for (int64_t i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int c = 0; c < j; ++c) acc += (1.0 * y[c]) * Z[(size_t)c * n + i];
      x[i] += acc;
    }
    if (res->final_residual <= tol) {
      sys_apply(s, k, x, w);
      for (int64_t i = 0; i < n; ++i) r[i] = b[i];
      for (int64_t i = 0; i < n; ++i) r[i] += -1.0 * w[i];
      res->final_residual = sqrt(fxo_dot(n, r, r, lanes));
      if (res->final_residual <= tol) {
        res->converged = 1;
        break;
      }
    }
    if (res->iterations >= max_it) break;
  }
#undef HIJ
  free(V);
  free(Z);
  free(w);
  free(r);
  free(inv);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(h);
  free(h2);
  free(y);
  return status;
}

/* PathTimeSeries::getFactor (ground_motion.cpp:43-74): the interval [time[l], time[l + 1]] is
 * found by walking from the last call's l, forward or back; a time equal to the walk's lower
 * end returns that sample exactly; outside the data, 0 */
static double path_time(const fxo_series* s, int32_t* loc, double t) {
  const double* v = s->values;
  const double* tm = s->time;
  const int32_t n = s->count;
  int32_t l = *loc;
  double t1 = tm[l];
  if (t < t1 && l == 0) return 0.0;
  if (t == t1) return s->factor * v[l];
  if (t > t1 && l == n - 1) return 0.0;
  double t2 = tm[l + 1];
  if (t > t2) {
    while (t > t2 && l < n - 2) {
      ++l;
      t1 = t2;
      t2 = tm[l + 1];
    }
    *loc = l;
    if (t > t2) return 0.0;
  } else if (t < t1) {
    while (t < t1 && l > 0) {
      --l;
      t2 = t1;
      t1 = tm[l];
    }
    *loc = l;
    if (t < t1) return 0.0;
  }
  return s->factor * (v[l] + (v[l + 1] - v[l]) * (t - t1) / (t2 - t1));
}

double fxo_series_value(const fxo_series* s, double fact, double t) {
  if (t < 0.0) return 0.0; /* GroundMotion::getAccel: 0 before time 0 */
  if (s->time) {
    int32_t l0 = 0;
    return fact * path_time(s, s->loc ? s->loc : &l0, t);
  }
  double a = 0.0;
  if (!(t < s->start)) {   /* PathSeries::getFactor */
    const double incr = (t - s->start) / s->dt;
    const long long incr1 = (long long)floor(incr);
    const long long incr2 = incr1 + 1;
    if (incr2 >= s->count) {
      a = s->use_last ? s->factor * s->values[s->count - 1] : 0.0;
    } else {
      const double value1 = s->values[incr1], value2 = s->values[incr2];
      a = s->factor * (value1 + (value2 - value1) * (incr - (double)incr1));
    }
  }
  return fact * a;
}

typedef struct nm_state {
  fxo_system* s;
  const fxo_newmark_opts* o;
  int ground;
  fxo_newmark_stats* st;
  int64_t n;
  int32_t step;
  double *u, *v, *a, *du, *ut, *vt, *at, *du0, *f, *tmp, *b, *mr, *scratch, *x0;
  fxo_series series; /* the ground series with the run's interval walk (time-array series) */
  int32_t loc;
  /* LoadControl (:262-269) and the kept tangent of ModifiedNewton -factoronce (:166-170) */
  double dlambda, dmin, dmax, spec_iter, last_iter, tangent_h;
  int kept_tangent;
} nm_state;

/* form_unbalance (:212-219) and form_residual (:55-68): B = -f(u, v) - M a - a_g M r; static
 * stages: the elements see v = 0 and there is no inertia */
static void nm_unbalance(nm_state* w, double ag) {
  const fxo_coeffs mass = {0.0, 0.0, 1.0, 0};
  const int is_static = w->o->analysis == 1;
  sys_force(w->s, w->u, is_static ? NULL : w->v, w->f);
  if (!is_static) sys_apply(w->s, &mass, w->a, w->tmp);
  for (int64_t q = 0; q < w->n; ++q) {
    double r = -w->f[q];
    if (!is_static) r -= w->tmp[q];
    if (w->ground >= 0) r -= ag * w->mr[q];
    w->b[q] = r;
  }
}

/* operator_symmetry (matrix_free_operator.cpp:5-16): 1 when the operator is SPD (ElasticIsotropic
 * SPD; PressureIndependMultiYield SPSD, made definite by a mass term) */
static int nm_spd(const fxo_system* s, const fxo_coeffs* k) {
  const int positive = k->stiffness >= 0.0 && k->damping >= 0.0 && k->mass >= 0.0 &&
                       (k->stiffness > 0.0 || k->mass > 0.0);
  if (!s->pimy) return positive;
  return positive && k->mass > 0.0;
}

/* KrylovSolver::solve (krylov.cpp:25-42): CG, GMRES(30) from the same initial guess when CG meets
 * p^T A p <= 0 (iterations added, indefinite set); GMRES from the start for a non-SPD operator */
static void nm_solve(nm_state* w, const fxo_coeffs* k, fxo_solve_result* r) {
  const fxo_newmark_opts* o = w->o;
  const int64_t n = w->n;
  if (!nm_spd(w->s, k)) {
    fxo_gmres_system(w->s, k, w->b, w->du, o->cg_rtol, 0.0, o->cg_max_it, 30, o->lanes, r);
    return;
  }
  memcpy(w->x0, w->du, (size_t)n * sizeof(double));
  fxo_cg_system(w->s, k, w->b, w->du, o->cg_rtol, 0.0, o->cg_max_it, o->lanes, r);
  if (!r->indefinite) return;
  const int32_t cg_it = r->iterations;
  memcpy(w->du, w->x0, (size_t)n * sizeof(double));
  fxo_gmres_system(w->s, k, w->b, w->du, o->cg_rtol, 0.0, o->cg_max_it, 30, o->lanes, r);
  r->iterations += cg_it;
  r->indefinite = 1;
}

/* one attempt of size h (:274-372); 1 = converged and committed */
static int nm_attempt(nm_state* w, double h) {
  const fxo_newmark_opts* o = w->o;
  const int64_t n = w->n;
  const double gamma = o->gamma, beta = o->beta;
  const int is_static = o->analysis == 1, linear = o->algorithm == 1, modified = o->algorithm == 2;
  const int factor_once = modified && o->factor_once, initial = modified && o->modified_initial;
  const double time0 = w->st->time;
  if (is_static) { /* LoadControl::newStep */
    w->dlambda *= w->spec_iter / w->last_iter;
    if (w->dlambda < w->dmin) w->dlambda = w->dmin;
    else if (w->dlambda > w->dmax) w->dlambda = w->dmax;
    memcpy(w->ut, w->u, (size_t)n * sizeof(double));
    w->st->time += w->dlambda;
    w->last_iter = 0;
  } else {
    const double a1 = 1.0 - gamma / beta, a2 = h * (1.0 - 0.5 * gamma / beta);
    const double a3 = -1.0 / (beta * h), a4 = 1.0 - 0.5 / beta;
    for (int64_t q = 0; q < n; ++q) { /* newmark_predict (:29-40) */
      w->ut[q] = w->u[q];
      w->vt[q] = w->v[q];
      w->at[q] = w->a[q];
      w->v[q] = w->v[q] * a1 + w->at[q] * a2;
      w->a[q] = w->a[q] * a4 + w->vt[q] * a3;
    }
    w->st->time += h;
  }
  memcpy(w->du0, w->du, (size_t)n * sizeof(double));
  double ag = 0.0;
  if (w->ground >= 0)
    ag = o->ground_series ? fxo_series_value(&w->series, o->ground_factor, w->st->time)
                          : (o->ground_accel ? o->ground_accel[w->step] : 0.0);
  fxo_coeffs k = {1.0, gamma / (beta * h), 1.0 / (beta * h * h), initial};
  if (is_static) k = (fxo_coeffs){1.0, 0.0, 0.0, initial};
  int iterations = 0, ok = 1;
  if (linear) { /* :301-313 */
    w->tangent_h = h;
    sys_store(w->s, w->u, w->scratch);
    nm_unbalance(w, ag);
    fxo_solve_result r;
    nm_solve(w, &k, &r);
    w->st->krylov_iterations += r.iterations;
    if (!r.converged) ok = 0;
    else
      for (int64_t q = 0; q < n; ++q) {
        w->u[q] += w->du[q];
        if (!is_static) {
          w->v[q] += w->du[q] * k.damping;
          w->a[q] += w->du[q] * k.mass;
        }
      }
    iterations = 1;
  } else {
    nm_unbalance(w, ag);
    if (modified && !(factor_once && w->kept_tangent && h == w->tangent_h)) { /* :324-328 */
      w->tangent_h = h;
      sys_store(w->s, w->u, w->scratch);
    }
    w->kept_tangent = factor_once;
    int iter = 1;
    for (;;) {
      if (!modified) { /* update_tangent: the tangents at u */
        w->tangent_h = h;
        sys_store(w->s, w->u, w->scratch);
      }
      fxo_solve_result r;
      nm_solve(w, &k, &r);
      w->st->krylov_iterations += r.iterations;
      if (!r.converged && !(r.final_residual <= 1e-6 * r.rhs_norm)) {
        ok = 0;
        break;
      }
      for (int64_t q = 0; q < n; ++q) { /* newmark_update (:43-52), or u += dU */
        w->u[q] += w->du[q];
        if (!is_static) {
          w->v[q] += w->du[q] * k.damping;
          w->a[q] += w->du[q] * k.mass;
        }
      }
      nm_unbalance(w, ag);
      double value;
      if (o->test == 2) value = 0.5 * fabs(fxo_dot(n, w->du, w->b, o->lanes));
      else if (o->test == 0) value = sqrt(fxo_dot(n, w->b, w->b, o->lanes));
      else value = sqrt(fxo_dot(n, w->du, w->du, o->lanes));
      iterations = iter;
      w->st->last_test = value;
      if (value <= o->tol) break;
      if (iter >= o->max_iter) {
        ok = 0;
        w->kept_tangent = 0; /* OpenSees: factorOnce = 1 after a failed test */
        break;
      }
      ++iter;
    }
  }
  w->st->newton_iterations += iterations;
  if (is_static) w->last_iter = iterations; /* LoadControl counts the failed attempt's updates too */
  if (!ok) {
    memcpy(w->u, w->ut, (size_t)n * sizeof(double));
    if (!is_static) {
      memcpy(w->v, w->vt, (size_t)n * sizeof(double));
      memcpy(w->a, w->at, (size_t)n * sizeof(double));
    }
    memcpy(w->du, w->du0, (size_t)n * sizeof(double));
    w->st->time = time0;
    ++w->st->failed_attempts;
    return 0;
  }
  if (iterations > w->st->max_newton_iterations) w->st->max_newton_iterations = iterations;
  ++w->st->commits;
  if (w->s->pimy) { /* commit_state */
    double* t = w->s->C;
    w->s->C = w->s->T;
    w->s->T = t;
  }
  return 1;
}

/* recover (:385-402); static stages retry only */
static int nm_recover(nm_state* w, double h, int level) {
  if (nm_attempt(w, h)) return 1;
  for (int r = 1; r <= w->o->max_retries; ++r) {
    ++w->st->retries;
    if (nm_attempt(w, h)) return 1;
  }
  if (w->o->analysis == 1 || level >= w->o->num_sublevels) return 0;
  ++w->st->substepped;
  const double hs = h / (double)w->o->num_substeps;
  for (int i = 0; i < w->o->num_substeps; ++i)
    if (!nm_recover(w, hs, level + 1)) return 0;
  return 1;
}

int fxo_newmark(fxo_system* s, const fxo_newmark_opts* o, int ground_direction, double* u, double* v, double* a,
                double* du, fxo_newmark_stats* st) {
  const int64_t n = s->c->neq;
  memset(st, 0, sizeof(*st));
  st->time = o->time;
  nm_state w;
  w.s = s;
  w.o = o;
  w.ground = ground_direction;
  w.st = st;
  w.n = n;
  w.step = 0;
  w.u = u;
  w.v = v;
  w.a = a;
  w.du = du;
  /* a time-array series walks on from the caller's interval (GroundMotions keeps it for the
   * pattern's life) or from interval 0, and the run keeps its walk from attempt to attempt */
  if (o->ground_series) {
    w.series = *o->ground_series;
    w.loc = w.series.loc ? *w.series.loc : 0;
    w.series.loc = &w.loc;
  }
  const int adapt = o->load_min_increment != 0.0 && o->load_max_increment != 0.0;
  w.dlambda = o->load_increment;
  w.dmin = adapt ? o->load_min_increment : o->load_increment;
  w.dmax = adapt ? o->load_max_increment : o->load_increment;
  w.spec_iter = adapt ? (double)o->load_num_iter : 1.0;
  w.last_iter = w.spec_iter;
  w.tangent_h = o->dt;
  w.kept_tangent = 0;
  double** bufs[] = {&w.ut, &w.vt, &w.at, &w.du0, &w.f, &w.tmp, &w.b, &w.scratch, &w.x0};
  for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i) *bufs[i] = (double*)malloc((size_t)(n + 1) * sizeof(double));
  w.mr = (double*)calloc((size_t)n + 1, sizeof(double));
  if (ground_direction >= 0) /* mass_times_ground */
    fxo_mass_times_ground(s->c, s->elastic, s->pimy, ground_direction, w.mr, s->nthreads);
  int status = 0;
  for (w.step = 0; w.step < o->nsteps; ++w.step) {
    if (!nm_recover(&w, o->dt, 0)) {
      status = 1;
      break;
    }
    st->steps = w.step + 1;
  }
  for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); ++i) free(*bufs[i]);
  free(w.mr);
  if (o->ground_series && o->ground_series->loc) *o->ground_series->loc = w.loc;
  return status;
}
