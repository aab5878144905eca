/*
 * fxo PressureIndependMultiYield (the J2 soil return map) — TEST INFRASTRUCTURE (the oracle).
 *
 * Restates src/physics/materials/pressure_independ_multi_yield.hpp (setup :110-139,
 * setUpSurfaces :275-381, initial_state :142-150, update :228-240, getStress :455-509, finish
 * :512-539, elast2Plast :542-571, yieldFunc :573-580, deviatorScaling :583-609,
 * initSurfaceUpdate :611-631, paramScaling :633-658, setTrialStress :660-666, setSubStrainRate
 * :668-692, getContactStress :703-712, isLoadReversal :714-722, getSurfaceNormal :724-730,
 * getLoadingFunc :732-750, stressCorrection :753-774, updateActiveSurface :776-832,
 * updateInnerSurface :834-849, isCrossingNextSurface :851-856, getTangent :860-886) with the
 * tensor helpers of src/physics/common/t2_vector.hpp (:32-220).
 *
 * State per point: NSTATE = 30 + 8 NS doubles in the reference's layout (STATE :74-85), so the
 * committed/trial buffers are interchangeable with the reference's. The operations keep the
 * reference's order (its goldens, 34 OpenSees strain paths, are matched at 1e-14 relative and
 * mostly bit for bit on a compiler that does not contract multiply-adds).
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "fxo.h"

#define NS FXO_PIMY_NS
#define LOW_LIMIT (20. * DBL_EPSILON)
#define UP_LIMIT 1.0e+30
#define BIG 1.0e200
/* state offsets */
#define S_STRESS 0
#define S_STRAIN 13
#define S_ACTIVE 26
#define S_E2P 27
#define S_G 28
#define S_K 29
#define S_SIZES 30
#define S_MODULI (30 + NS)
#define S_CENTERS (30 + 2 * NS)

/* ---- 6-vectors (OpenSees Vector semantics) --------------------------------------------- */
static void v6copy(double* a, const double* b) { for (int i = 0; i < 6; ++i) a[i] = b[i]; }
static void v6zero(double* a) { for (int i = 0; i < 6; ++i) a[i] = 0.0; }
static void v6sub(double* a, const double* b) { for (int i = 0; i < 6; ++i) a[i] -= b[i]; }
static void v6add(double* a, const double* b) { for (int i = 0; i < 6; ++i) a[i] += b[i]; }
static void v6scale(double* a, double f) { for (int i = 0; i < 6; ++i) a[i] *= f; }
static void v6div(double* a, double f) {
  if (f == 0.0) {
    for (int i = 0; i < 6; ++i) a[i] = BIG;
  } else {
    for (int i = 0; i < 6; ++i) a[i] /= f;
  }
}
/* a = thisFact * a + otherFact * other, with Vector::addVector's special cases */
static void v6axpby(double* a, double thisFact, const double* other, double otherFact) {
  if (otherFact == 0.0) {
    if (thisFact == 1.0) return;
    if (thisFact == 0.0) {
      v6zero(a);
      return;
    }
    for (int i = 0; i < 6; ++i) a[i] *= thisFact;
    return;
  }
  if (thisFact == 1.0) {
    if (otherFact == 1.0) for (int i = 0; i < 6; ++i) a[i] += other[i];
    else if (otherFact == -1.0) for (int i = 0; i < 6; ++i) a[i] -= other[i];
    else for (int i = 0; i < 6; ++i) a[i] += other[i] * otherFact;
  } else if (thisFact == 0.0) {
    if (otherFact == 1.0) for (int i = 0; i < 6; ++i) a[i] = other[i];
    else if (otherFact == -1.0) for (int i = 0; i < 6; ++i) a[i] = -(other[i]);
    else for (int i = 0; i < 6; ++i) a[i] = other[i] * otherFact;
  } else {
    if (otherFact == 1.0) for (int i = 0; i < 6; ++i) a[i] = a[i] * thisFact + other[i];
    else if (otherFact == -1.0) for (int i = 0; i < 6; ++i) a[i] = a[i] * thisFact - other[i];
    else for (int i = 0; i < 6; ++i) a[i] = a[i] * thisFact + other[i] * otherFact;
  }
}
/* tensor double-dot in vector form (shear counted twice) */
static double ddot(const double* a, const double* b) {
  double r = 0.;
  for (int i = 0; i < 3; ++i) r += a[i] * b[i] + 2 * a[i + 3] * b[i + 3];
  return r;
}

/* ---- T2: the tensor, its deviator and its mean ------------------------------------------ */
typedef struct { double t2[6], dev[6], vol; } t2v;
static void t2_set(t2v* s, const double* init, int engr) {
  v6copy(s->t2, init);
  s->vol = (s->t2[0] + s->t2[1] + s->t2[2]) / 3.0;
  for (int i = 0; i < 3; ++i) {
    s->dev[i] = s->t2[i] - s->vol;
    s->dev[i + 3] = s->t2[i + 3];
    if (engr == 1) {
      s->dev[i + 3] /= 2.;
      s->t2[i + 3] /= 2.;
    }
  }
}
static void t2_set_dev(t2v* s, const double* deviat, double volume) {
  s->vol = volume;
  const double dv = (deviat[0] + deviat[1] + deviat[2]) / 3.;
  for (int i = 0; i < 3; ++i) {
    s->dev[i] = deviat[i] - dv;
    s->dev[i + 3] = deviat[i + 3];
    s->t2[i] = s->dev[i] + s->vol;
    s->t2[i + 3] = s->dev[i + 3];
  }
}
static void t2_engr(const t2v* s, double* out) {
  v6copy(out, s->t2);
  for (int i = 0; i < 3; ++i) out[i + 3] *= 2.;
}
static double t2_dev_len(const t2v* s) { return sqrt(ddot(s->dev, s->dev)); }
static int t2_is_zero(const t2v* s) {
  for (int i = 0; i < 6; ++i)
    if (s->t2[i] != 0.0) return 0;
  return 1;
}
static void t2_load(t2v* s, const double* st, int k) {
  for (int i = 0; i < 6; ++i) s->t2[i] = st[k + i];
  for (int i = 0; i < 6; ++i) s->dev[i] = st[k + 6 + i];
  s->vol = st[k + 12];
}
static void t2_store(const t2v* s, double* st, int k) {
  for (int i = 0; i < 6; ++i) st[k + i] = s->t2[i];
  for (int i = 0; i < 6; ++i) st[k + 6 + i] = s->dev[i];
  st[k + 12] = s->vol;
}

/* MultiYieldSurface secondOrderEqn: the smaller non-negative root (t2_vector.hpp:181-220) */
static double quad_root(double A, double B, double C, int* ok) {
  if (A == 0) {
    *ok = 0;
    return NAN;
  }
  if (C == 0) return 0;
  if (B == 0) {
    if (C / A > 0) {
      *ok = 0;
      return NAN;
    }
    return sqrt(-C / A);
  }
  const double det = B * B - 4. * A * C;
  if (det < 0) {
    *ok = 0;
    return NAN;
  }
  double r1 = B > 0 ? (-B - sqrt(det)) / (2. * A) : (-B + sqrt(det)) / (2. * A);
  double r2 = C / (A * r1);
  if (r1 < 0 && r2 < 0) {
    if (fabs(r1) < LOW_LIMIT) r1 = 0.;
    else if (fabs(r2) < LOW_LIMIT) r2 = 0.;
  }
  if (r1 < 0 && r2 < 0) {
    *ok = 0;
    return NAN;
  }
  if (r1 < 0) return r2;
  if (r2 < 0) return r1;
  return r1 > r2 ? r2 : r1;
}

static void elastic_D(double G, double K, double* D) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double v = 0.;
      if (i == j) v += G;
      if (i < 3 && j < 3 && i == j) v += G;
      if (i < 3 && j < 3) v += (K - 2. * G / 3.);
      D[6 * i + j] = v;
    }
}

/* ---- setup ----------------------------------------------------------------------------- */
int fxo_pimy_setup(const double* in, int nsurf, const double* gredu, fxo_pimy_params* p) {
  double rho = in[0], G = in[1], K = in[2], cohesion = in[3], gamma_peak = in[4], phi = in[5], pref = in[6],
         pdc = in[7];
  if (G <= 0 || K <= 0) return 5;
  if (phi < 0.) phi = 0.;
  if (phi == 0. && cohesion <= 0.) return 5;
  if (cohesion <= 0) cohesion = 0.;
  if (gamma_peak <= 0 || pref <= 0) return 5;
  if (pdc < 0) pdc = 0.;
  if (pdc > 0 && phi == 0) pdc = 0.;
  if (nsurf <= 0) nsurf = 10;
  if (nsurf > 100) nsurf = 100;
  if (rho < 0) rho = 0.;
  if (nsurf > NS) return 1;
  if (gredu && nsurf < 2) return 4;
  memset(p, 0, sizeof(*p));
  p->load_stage = 0;
  p->nsurf = nsurf;
  p->rho = rho;
  p->G = G;
  p->K = K;
  p->phi = phi;
  p->gamma_peak = gamma_peak;
  p->pref = -pref; /* compression negative */
  p->cohesion = cohesion;
  p->pdc = pdc;
  p->residual_press = 0.;
  const double pi = 3.14159265358979;
  double residual = 0., cone = 0., peak = 0., ref_strain = 0.;
  if (!gredu) {
    if (phi > 0) {
      const double sinPhi = sin(phi * pi / 180.);
      const double Mnys = 6. * sinPhi / (3. - sinPhi);
      residual = 3. * cohesion / (sqrt(2.) * Mnys);
      cone = -(p->pref - residual);
      peak = sqrt(2.) * cone * Mnys / 3.;
      ref_strain = (gamma_peak * peak) / (G * gamma_peak - peak);
    } else {
      peak = 2 * sqrt(2.) * cohesion / 3;
      ref_strain = (gamma_peak * peak) / (G * gamma_peak - peak);
      residual = 0.;
    }
    const double inc = peak / nsurf;
    for (int ii = 1; ii <= nsurf; ++ii) {
      const double s1 = ii * inc, s2 = s1 + inc;
      const double e1 = s1 * ref_strain / (G * ref_strain - s1);
      const double e2 = s2 * ref_strain / (G * ref_strain - s2);
      const double size = phi > 0. ? 3. * s1 / sqrt(2.) / cone : 3. * s1 / sqrt(2.);
      const double epm = 2. * (s2 - s1) / (e2 - e1);
      double pm = (2. * G - epm) <= 0 ? UP_LIMIT : (2. * G * epm) / (2. * G - epm);
      if (pm < 0) pm = 0;
      if (pm > UP_LIMIT) pm = UP_LIMIT;
      if (ii == nsurf) pm = 0;
      p->size[ii] = size;
      p->modulus[ii] = pm;
    }
  } else {
    const int last = 2 * (nsurf - 1);
    const double tmax = G * gredu[last] * gredu[last + 1];
    if (phi > 0) {
      const double Mnys = -(sqrt(3.) * tmax - 2. * cohesion) / p->pref;
      if (Mnys <= 0) {
        cohesion = sqrt(3.) / 2 * tmax;
        phi = 0.;
        cone = 1.;
        residual = 0.;
      } else {
        const double sinPhi = 3 * Mnys / (6 + Mnys);
        if (sinPhi < 0. || sinPhi > 1.) return 2;
        residual = 2. * cohesion / Mnys;
        if (residual < 0.01 * p->pref) residual = 0.01 * p->pref;
        cone = -(p->pref - residual);
        phi = asin(sinPhi) * 180 / pi;
      }
    } else {
      cohesion = sqrt(3.) / 2 * tmax;
      cone = 1.;
      residual = 0.;
    }
    for (int i = 1; i < nsurf; ++i) {
      const int ii = 2 * (i - 1);
      const double e1 = gredu[ii], s1 = G * gredu[ii + 1] * e1;
      const double e2 = gredu[ii + 2], s2 = G * gredu[ii + 3] * e2;
      const double size = sqrt(3.) * s1 / cone;
      const double epm = 2. * (s2 - s1) / (e2 - e1);
      double pm = (2. * G - epm) <= 0 ? UP_LIMIT : (2. * G * epm) / (2. * G - epm);
      if (pm <= 0) return 3;
      if (pm > UP_LIMIT) pm = UP_LIMIT;
      p->size[i] = size;
      p->modulus[i] = pm;
      if (i == nsurf - 1) {
        p->size[i + 1] = sqrt(3.) * s2 / cone;
        p->modulus[i + 1] = 0;
      }
    }
  }
  p->residual_press = residual;
  p->phi = phi;
  p->cohesion = cohesion;
  return 0;
}

int fxo_pimy_nstate(void) { return 30 + 8 * NS; }

void fxo_pimy_initial_state(const fxo_pimy_params* p, double* s) {
  for (int k = 0; k < 30 + 8 * NS; ++k) s[k] = 0.0;
  s[S_G] = p->G;
  s[S_K] = p->K;
  for (int i = 1; i <= NS; ++i) {
    s[S_SIZES + i - 1] = p->size[i];
    s[S_MODULI + i - 1] = p->modulus[i];
  }
}

/* ---- one point's update ----------------------------------------------------------------- */
typedef struct {
  const fxo_pimy_params* p;
  const double* C;
  double* T;
  int ok;
  t2v cur_stress, cur_strain, trial, rate, sub_rate;
  int committed_active, active, e2p;
  double G, K;
} point;

static double surf_size(const point* w, int i) { return i == 0 ? 0. : w->T[S_SIZES + i - 1]; }
static double surf_mod(const point* w, int i) { return i == 0 ? 0. : w->T[S_MODULI + i - 1]; }
static void surf_center(const point* w, int i, double* c) {
  if (i == 0) {
    v6zero(c);
    return;
  }
  for (int k = 0; k < 6; ++k) c[k] = w->T[S_CENTERS + 6 * (i - 1) + k];
}
static void surf_set_center(point* w, int i, const double* c) {
  for (int k = 0; k < 6; ++k) w->T[S_CENTERS + 6 * (i - 1) + k] = c[k];
}

static double yield_f(const point* w, const t2v* s, int i) {
  double tmp[6], c[6];
  v6copy(tmp, s->dev);
  surf_center(w, i, c);
  v6sub(tmp, c);
  const double sz = surf_size(w, i);
  return 3. / 2. * ddot(tmp, tmp) - sz * sz;
}

static void deviator_scaling(point* w, t2v* s, int i) {
  const int n = w->p->nsurf;
  for (int iter = 0; iter < 100000; ++iter) {
    const double diff = yield_f(w, s, i);
    if (i < n && diff < 0.) {
      const double sz = surf_size(w, i);
      const double dsz = sqrt(sz * sz + diff);
      double devia[6], tmp[6], c[6];
      v6copy(devia, s->dev);
      surf_center(w, i, c);
      for (int k = 0; k < 6; ++k) tmp[k] = devia[k] - c[k];
      double coeff = (sz - dsz) / dsz;
      if (coeff < 1.e-13) coeff = 1.e-13;
      v6axpby(devia, 1.0, tmp, coeff);
      t2_set_dev(s, devia, s->vol);
      continue;
    }
    if (i == n && fabs(diff) > LOW_LIMIT) {
      const double sz = surf_size(w, i);
      double nd[6] = {0, 0, 0, 0, 0, 0};
      v6axpby(nd, 0.0, s->dev, sz / sqrt(diff + sz * sz));
      t2_set_dev(s, nd, s->vol);
    }
    return;
  }
  w->ok = 0;
}

static void init_surface_update(point* w) {
  if (w->committed_active == 0) return;
  const int n = w->p->nsurf;
  double devia[6];
  v6copy(devia, w->cur_stress.dev);
  const double Ms = sqrt(3. / 2. * ddot(devia, devia));
  if (w->committed_active < n) {
    double nc[6] = {0, 0, 0, 0, 0, 0};
    v6axpby(nc, 0.0, devia, 1.0 - surf_size(w, w->committed_active) / Ms);
    surf_set_center(w, w->committed_active, nc);
  }
  for (int i = 1; i < w->committed_active; ++i) {
    double nc[6];
    const double f = 1. - surf_size(w, i) / Ms;
    for (int c = 0; c < 6; ++c) nc[c] = devia[c] * f;
    surf_set_center(w, i, nc);
  }
}

static void param_scaling(point* w) {
  const fxo_pimy_params* p = w->p;
  if (p->phi == 0.) return;
  const double conHeig = -(w->cur_stress.vol - p->residual_press);
  double scale = -conHeig / (p->pref - p->residual_press);
  scale = pow(scale, p->pdc);
  w->G *= scale;
  w->K *= scale;
  for (int i = 1; i <= p->nsurf; ++i) {
    const double pm = surf_mod(w, i) * scale;
    const double sz = surf_size(w, i) * conHeig;
    w->T[S_SIZES + i - 1] = sz;
    w->T[S_MODULI + i - 1] = pm;
    const double zero[6] = {0, 0, 0, 0, 0, 0};
    surf_set_center(w, i, zero);
  }
}

static void elast_to_plast(point* w) {
  const fxo_pimy_params* p = w->p;
  if (p->load_stage != 1 || w->e2p == 1) return;
  w->e2p = 1;
  if (w->cur_stress.vol > 0. && p->phi > 0.) {
    double d[6];
    v6copy(d, w->cur_stress.dev);
    t2_set(&w->cur_stress, d, 0);
  }
  param_scaling(w);
  if (t2_dev_len(&w->cur_stress) == 0.) return;
  while (yield_f(w, &w->cur_stress, ++w->committed_active) > 0) {
    if (w->committed_active == p->nsurf) {
      deviator_scaling(w, &w->cur_stress, p->nsurf);
      init_surface_update(w);
      return;
    }
  }
  w->committed_active--;
  init_surface_update(w);
}

static void set_trial_stress(point* w, const t2v* in) {
  const t2v s = *in;
  double devia[6];
  v6copy(devia, s.dev);
  v6axpby(devia, 1.0, w->sub_rate.dev, 2. * w->G);
  t2_set_dev(&w->trial, devia, s.vol);
}

static int to_int(double x) {
  if (!(x == x)) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return -2147483647 - 1;
  return (int)x;
}

static int set_sub_strain_rate(point* w) {
  const int n = w->p->nsurf;
  if (t2_is_zero(&w->rate)) return 0;
  double epm;
  if (w->active == 0) epm = 2 * w->G;
  else {
    const double pm = surf_mod(w, w->active);
    epm = 2 * w->G * pm / (2 * w->G + pm);
  }
  double inc[6] = {0, 0, 0, 0, 0, 0};
  v6axpby(inc, 0.0, w->rate.dev, epm);
  t2v s;
  t2_set(&s, inc, 0);
  const double single = surf_size(w, n) / n;
  const double total = 3. * (sqrt(1. / 3.) * t2_dev_len(&s)) / sqrt(2.);
  int m = to_int(total / single + 1);
  if (m > n) m = n;
  v6copy(inc, w->rate.t2);
  v6div(inc, m);
  t2_set(&w->sub_rate, inc, 0);
  return m;
}

static void surface_normal(const point* w, const t2v* s, double* nrm) {
  double c[6];
  v6copy(nrm, s->dev);
  surf_center(w, w->active, c);
  v6sub(nrm, c);
  v6div(nrm, sqrt(ddot(nrm, nrm)));
}

static int load_reversal(const point* w) {
  if (w->active == 0) return 0;
  double nrm[6], a[6];
  surface_normal(w, &w->cur_stress, nrm);
  v6copy(a, w->trial.dev);
  v6sub(a, w->cur_stress.dev);
  return ddot(a, nrm) < 0 ? 1 : 0;
}

static void contact_stress(const point* w, t2v* cs) {
  double c[6], devia[6];
  surf_center(w, w->active, c);
  v6copy(devia, w->trial.dev);
  v6sub(devia, c);
  const double Ms = sqrt(3. / 2. * ddot(devia, devia));
  v6scale(devia, surf_size(w, w->active) / Ms);
  v6add(devia, c);
  t2_set_dev(cs, devia, w->trial.vol);
}

static double loading_func(const point* w, const t2v* cs, const double* nrm, int crossed) {
  const double t1 = 2. * w->G, t2 = surf_mod(w, w->active);
  const double t = t1 + t2;
  double tmp[6];
  v6copy(tmp, w->trial.dev);
  v6sub(tmp, cs->dev);
  double lf = ddot(nrm, tmp) / t;
  if (crossed) {
    const double t3 = surf_mod(w, w->active - 1);
    lf *= (t3 - t2) / t3;
  }
  return lf;
}

static int crossing_next(const point* w) {
  if (w->active == w->p->nsurf) return 0;
  return yield_f(w, &w->trial, w->active + 1) > 0 ? 1 : 0;
}

static void stress_correction(point* w, int crossed) {
  for (;;) {
    t2v cs;
    contact_stress(w, &cs);
    double nrm[6];
    surface_normal(w, &cs, nrm);
    const double lf = loading_func(w, &cs, nrm, crossed);
    double devia[6] = {0, 0, 0, 0, 0, 0};
    v6axpby(devia, 0.0, nrm, -2 * w->G * lf);
    v6add(devia, w->trial.dev);
    t2_set_dev(&w->trial, devia, w->trial.vol);
    deviator_scaling(w, &w->trial, w->active);
    if (crossing_next(w)) {
      w->active++;
      crossed = 1;
      continue;
    }
    return;
  }
}

static void update_active_surface(point* w) {
  const int n = w->p->nsurf;
  if (w->active == n) return;
  double t1[6], t2[6], tmp[6], cen[6], outc[6];
  surf_center(w, w->active, cen);
  const double sz = surf_size(w, w->active);
  surf_center(w, w->active + 1, outc);
  const double outsz = surf_size(w, w->active + 1);
  v6copy(t1, w->trial.dev);
  v6sub(t1, cen);
  v6copy(t2, cen);
  v6sub(t2, outc);
  double A = ddot(t1, t1);
  double B = 2. * ddot(t1, t2);
  double Cc = ddot(t2, t2) - 2. / 3. * outsz * outsz;
  double X = quad_root(A, B, Cc, &w->ok);
  if (fabs(X - 1.) < LOW_LIMIT) X = 1.;
  if (X < 1.) {
    w->ok = 0;
    return;
  }
  v6copy(tmp, cen);
  v6axpby(tmp, 1.0, t1, X);
  v6scale(tmp, (1.0 - sz / outsz));
  v6copy(t2, cen);
  v6axpby(t2, 1.0, outc, -sz / outsz);
  v6sub(tmp, t2);
  t2v dir;
  t2_set(&dir, tmp, 0);
  if (t2_dev_len(&dir) < LOW_LIMIT) return;
  v6copy(tmp, dir.dev);
  A = ddot(tmp, tmp);
  B = -2 * ddot(t1, tmp);
  if (fabs(B) < LOW_LIMIT) B = 0.;
  Cc = ddot(t1, t1) - 2. / 3. * sz * sz;
  if (fabs(Cc) < LOW_LIMIT || fabs(Cc) / ddot(t1, t1) < LOW_LIMIT) return;
  if ((fabs(B) < 1.0e-10) && (fabs(Cc) < 1.0e-10)) return;
  if (B > 0. || Cc < 0.) {
    w->ok = 0;
    return;
  }
  X = quad_root(A, B, Cc, &w->ok);
  v6axpby(cen, 1.0, tmp, X);
  surf_set_center(w, w->active, cen);
}

static void update_inner_surface(point* w) {
  if (w->active <= 1) return;
  double devia[6], cen[6];
  v6copy(devia, w->cur_stress.dev);
  surf_center(w, w->active, cen);
  const double sz = surf_size(w, w->active);
  for (int i = 1; i < w->active; ++i) {
    double nc[6];
    v6copy(nc, cen);
    v6sub(nc, devia);
    v6scale(nc, surf_size(w, i) / sz);
    v6add(nc, devia);
    surf_set_center(w, i, nc);
  }
}

static void tangent(const point* w, double* D) {
  if (w->p->load_stage != 1) {
    elastic_D(w->G, w->K, D);
    return;
  }
  double coeff, devia[6] = {0, 0, 0, 0, 0, 0};
  if (w->active > 0) {
    double c[6];
    v6copy(devia, w->trial.dev);
    surf_center(w, w->active, c);
    v6sub(devia, c);
    const double sz = surf_size(w, w->active), pm = surf_mod(w, w->active);
    coeff = 6. * w->G * w->G / (2. * w->G + pm) / sz / sz;
  } else {
    coeff = 0.;
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double v = -coeff * devia[i] * devia[j];
      if (i == j) v += w->G;
      if (i < 3 && j < 3 && i == j) v += w->G;
      if (i < 3 && j < 3) v += (w->K - 2. * w->G / 3.);
      D[6 * i + j] = v;
    }
}

void fxo_pimy_committed_tangent(const fxo_pimy_params* p, const double* C, double D[36]) {
  point w;
  memset(&w, 0, sizeof w);
  w.p = p;
  w.C = C;
  w.T = (double*)C; /* the surfaces are read from the committed buffer; nothing is written */
  t2_load(&w.trial, C, S_STRESS);
  w.active = (int)C[S_ACTIVE];
  w.G = C[S_G];
  w.K = C[S_K];
  tangent(&w, D);
}

void fxo_pimy_update(const fxo_pimy_params* p, const double* C, double* T, const double eps[6], double sig[6],
                     double D[36]) {
  point w;
  memset(&w, 0, sizeof w);
  w.p = p;
  w.C = C;
  w.T = T;
  w.ok = 1;
  /* committed state */
  t2_load(&w.cur_stress, C, S_STRESS);
  t2_load(&w.cur_strain, C, S_STRAIN);
  w.committed_active = (int)C[S_ACTIVE];
  w.active = w.committed_active;
  w.e2p = (int)C[S_E2P];
  w.G = C[S_G];
  w.K = C[S_K];
  /* strain increment (engineering shear halved) */
  {
    double d[6], cs[6];
    v6copy(d, eps);
    t2_engr(&w.cur_strain, cs);
    v6sub(d, cs);
    t2_set(&w.rate, d, 1);
  }
  /* committed surfaces into the trial buffer */
  for (int s = 0; s < NS; ++s) {
    T[S_SIZES + s] = C[S_SIZES + s];
    T[S_MODULI + s] = C[S_MODULI + s];
    for (int c = 0; c < 6; ++c) T[S_CENTERS + 6 * s + c] = C[S_CENTERS + 6 * s + c];
  }
  if (p->load_stage == 1 && w.e2p == 0) elast_to_plast(&w);
  if (p->load_stage != 1) {
    double De[36], a[6], r[6];
    tangent(&w, De);
    v6copy(a, w.cur_stress.t2);
    t2_engr(&w.rate, r);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) a[j] += De[6 * j + i] * r[i];
    t2_set(&w.trial, a, 0);
  } else {
    w.active = w.committed_active;
    w.sub_rate = w.rate;
    set_trial_stress(&w, &w.cur_stress);
    if (load_reversal(&w)) {
      update_inner_surface(&w);
      w.active = 0;
    }
    const int nsub = set_sub_strain_rate(&w);
    for (int i = 0; i < nsub; ++i) {
      if (i == 0) set_trial_stress(&w, &w.cur_stress);
      else set_trial_stress(&w, &w.trial);
      if (w.active == 0 && !crossing_next(&w)) continue;
      if (w.active == 0) w.active++;
      stress_correction(&w, 0);
      update_active_surface(&w);
    }
    double vol = w.K * (w.rate.vol * 3.);
    vol += w.cur_stress.vol;
    double d[6];
    v6copy(d, w.trial.dev);
    t2_set_dev(&w.trial, d, vol);
  }
  /* finish: tangent, outputs, trial buffer */
  tangent(&w, D);
  for (int i = 0; i < 6; ++i) sig[i] = w.ok ? w.trial.t2[i] : NAN;
  double sc[6];
  v6copy(sc, w.cur_strain.t2);
  v6add(sc, w.rate.t2);
  t2v strain_commit;
  t2_set(&strain_commit, sc, 0);
  t2_store(&w.trial, T, S_STRESS);
  t2_store(&strain_commit, T, S_STRAIN);
  T[S_E2P] = w.e2p;
  T[S_G] = w.G;
  T[S_K] = w.K;
  T[S_ACTIVE] = p->load_stage == 1 ? w.active : w.committed_active;
}
