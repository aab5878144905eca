// The PressureIndependMultiYield material point as device code (the algorithm and its sources
// are listed in fx_pimy.hpp). Every translation unit that includes this header starts with
// `#pragma clang fp contract(off)`, so that the yield tests and sub-increment counts see the
// same roundings as the oracle.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "fx_pimy.hpp"

namespace fx {

namespace pimy {

constexpr double kLow = 20.0 * DBL_EPSILON;
constexpr double kBig = 1.0e200;

struct V6 {
  double v[6];
};

__device__ inline V6 zero6() { return V6{{0, 0, 0, 0, 0, 0}}; }
// tensor double contraction in vector form (shear terms count twice)
__device__ inline double dd(const V6& a, const V6& b) {
  double r = 0.;
  for (int i = 0; i < 3; ++i) r += a.v[i] * b.v[i] + 2 * a.v[i + 3] * b.v[i + 3];
  return r;
}
// a = s * a + f * o with OpenSees Vector::addVector's special cases (t2_vector.hpp:40-75)
__device__ inline void axpby(V6& a, double s, const V6& o, double f) {
  if (f == 0.0) {
    if (s == 1.0) return;
    if (s == 0.0) {
      a = zero6();
      return;
    }
    for (int i = 0; i < 6; ++i) a.v[i] *= s;
    return;
  }
  for (int i = 0; i < 6; ++i) {
    if (s == 1.0) {
      if (f == 1.0) a.v[i] += o.v[i];
      else if (f == -1.0) a.v[i] -= o.v[i];
      else a.v[i] += o.v[i] * f;
    } else if (s == 0.0) {
      if (f == 1.0) a.v[i] = o.v[i];
      else if (f == -1.0) a.v[i] = -(o.v[i]);
      else a.v[i] = o.v[i] * f;
    } else {
      if (f == 1.0) a.v[i] = a.v[i] * s + o.v[i];
      else if (f == -1.0) a.v[i] = a.v[i] * s - o.v[i];
      else a.v[i] = a.v[i] * s + o.v[i] * f;
    }
  }
}
// a /= f, with the 1e200 fill of a division by zero (t2_vector.hpp:86-92)
__device__ inline void divide(V6& a, double f) {
  for (int i = 0; i < 6; ++i) a.v[i] = f == 0.0 ? kBig : a.v[i] / f;
}

// a symmetric tensor with its deviator and mean (soil::T2)
struct T2 {
  V6 t2, dev;
  double vol;
  __device__ void set(const V6& x, bool engineering) {
    t2 = x;
    vol = (t2.v[0] + t2.v[1] + t2.v[2]) / 3.0;
    for (int i = 0; i < 3; ++i) {
      dev.v[i] = t2.v[i] - vol;
      dev.v[i + 3] = t2.v[i + 3];
      if (engineering) {
        dev.v[i + 3] /= 2.;
        t2.v[i + 3] /= 2.;
      }
    }
  }
  __device__ void set_dev(const V6& d, double volume) {
    vol = volume;
    const double dv = (d.v[0] + d.v[1] + d.v[2]) / 3.;
    for (int i = 0; i < 3; ++i) {
      dev.v[i] = d.v[i] - dv;
      dev.v[i + 3] = d.v[i + 3];
      t2.v[i] = dev.v[i] + vol;
      t2.v[i + 3] = dev.v[i + 3];
    }
  }
  __device__ V6 engineering() const {
    V6 o = t2;
    for (int i = 0; i < 3; ++i) o.v[i + 3] *= 2.;
    return o;
  }
  __device__ double dev_len() const { return sqrt(dd(dev, dev)); }
  __device__ bool is_zero() const {
    for (int i = 0; i < 6; ++i)
      if (t2.v[i] != 0.0) return false;
    return true;
  }
};

// the smaller non-negative root of A x^2 + B x + C (MultiYieldSurface secondOrderEqn)
__device__ inline double quad_root(double A, double B, double C, bool& ok) {
  if (A == 0) {
    ok = false;
    return NAN;
  }
  if (C == 0) return 0;
  if (B == 0) {
    if (C / A > 0) {
      ok = false;
      return NAN;
    }
    return sqrt(-C / A);
  }
  const double det = B * B - 4. * A * C;
  if (det < 0) {
    ok = false;
    return NAN;
  }
  double r1 = B > 0 ? (-B - sqrt(det)) / (2. * A) : (-B + sqrt(det)) / (2. * A);
  double r2 = C / (A * r1);
  if (r1 < 0 && r2 < 0) {
    if (fabs(r1) < kLow) r1 = 0.;
    else if (fabs(r2) < kLow) r2 = 0.;
  }
  if (r1 < 0 && r2 < 0) {
    ok = false;
    return NAN;
  }
  if (r1 < 0) return r2;
  if (r2 < 0) return r1;
  return r1 > r2 ? r2 : r1;
}

__device__ inline int to_int(double x) {
  if (!(x == x)) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return -2147483647 - 1;
  return static_cast<int>(x);
}

// One material point: committed history C, trial history T (strided views into the SoA state).
struct Point {
  const PimyParams& p;
  const double* C;
  double* T;
  int64_t s;  // stride between fields
  bool ok = true;
  T2 cur_stress, cur_strain, trial, rate, sub_rate;
  int committed_active = 0, active = 0, e2p = 0;
  double G = 0, K = 0;
  // The surfaces whose centres get_stress copies from C into T: a batch keeps per point the
  // highest surface whose centre any update has written, and every centre above it holds its
  // initial value in both buffers, so the copy can stop there (ncopy) with T equal to C after
  // it, as if all were copied. moved: the highest surface this update writes.
  int ncopy;
  mutable int moved = 0;

  __device__ Point(const PimyParams& pp, const double* c, double* t, int64_t stride, int copy = -1)
      : p(pp), C(c), T(t), s(stride), ncopy(copy < 0 || copy > pp.nsurf ? pp.nsurf : copy) {}

  __device__ double c(int k) const { return C[k * s]; }
  __device__ double& t(int k) const { return T[k * s]; }
  // With phi = 0 the surfaces' sizes and moduli never change (paramScaling is a no-op,
  // :633-639): they are read from the parameter set and their history fields are left as
  // initialised (same values), so an update moves 40 doubles less per point.
  __device__ bool fixed_surfaces() const { return p.phi == 0.; }
  __device__ double size(int i) const {
    return i == 0 ? 0. : fixed_surfaces() ? p.size[i] : t(kPSizes + i - 1);
  }
  __device__ double modulus(int i) const {
    return i == 0 ? 0. : fixed_surfaces() ? p.modulus[i] : t(kPModuli + i - 1);
  }
  __device__ V6 center(int i) const {
    V6 o = zero6();
    if (i == 0) return o;
    for (int k = 0; k < 6; ++k) o.v[k] = t(kPCenters + 6 * (i - 1) + k);
    return o;
  }
  __device__ void set_center(int i, const V6& o) const {
    moved = i > moved ? i : moved;
    for (int k = 0; k < 6; ++k) t(kPCenters + 6 * (i - 1) + k) = o.v[k];
  }
  __device__ T2 load_t2(int k) const {
    T2 x;
    for (int i = 0; i < 6; ++i) x.t2.v[i] = c(k + i);
    for (int i = 0; i < 6; ++i) x.dev.v[i] = c(k + 6 + i);
    x.vol = c(k + 12);
    return x;
  }
  __device__ void store_t2(const T2& x, int k) const {
    for (int i = 0; i < 6; ++i) t(k + i) = x.t2.v[i];
    for (int i = 0; i < 6; ++i) t(k + 6 + i) = x.dev.v[i];
    t(k + 12) = x.vol;
  }

  __device__ double yield(const T2& x, int i) const {
    V6 d = x.dev;
    const V6 cc = center(i);
    for (int k = 0; k < 6; ++k) d.v[k] -= cc.v[k];
    const double sz = size(i);
    return 3. / 2. * dd(d, d) - sz * sz;
  }

  __device__ void deviator_scaling(T2& x, int i) {
    const int n = p.nsurf;
    for (int iter = 0; iter < 100000; ++iter) {
      const double diff = yield(x, i);
      if (i < n && diff < 0.) {
        const double sz = size(i);
        const double dsz = sqrt(sz * sz + diff);
        V6 devia = x.dev, tmp;
        const V6 cc = center(i);
        for (int k = 0; k < 6; ++k) tmp.v[k] = devia.v[k] - cc.v[k];
        double coeff = (sz - dsz) / dsz;
        if (coeff < 1.e-13) coeff = 1.e-13;
        axpby(devia, 1.0, tmp, coeff);
        x.set_dev(devia, x.vol);
        continue;
      }
      if (i == n && fabs(diff) > kLow) {
        const double sz = size(i);
        V6 nd = zero6();
        axpby(nd, 0.0, x.dev, sz / sqrt(diff + sz * sz));
        x.set_dev(nd, x.vol);
      }
      return;
    }
    ok = false;
  }

  __device__ void init_surface_update() {
    if (committed_active == 0) return;
    const int n = p.nsurf;
    const V6 devia = cur_stress.dev;
    const double Ms = sqrt(3. / 2. * dd(devia, devia));
    if (committed_active < n) {
      V6 nc = zero6();
      axpby(nc, 0.0, devia, 1.0 - size(committed_active) / Ms);
      set_center(committed_active, nc);
    }
    for (int i = 1; i < committed_active; ++i) {
      V6 nc;
      const double f = 1. - size(i) / Ms;
      for (int k = 0; k < 6; ++k) nc.v[k] = devia.v[k] * f;
      set_center(i, nc);
    }
  }

  __device__ void param_scaling() {
    if (p.phi == 0.) return;
    const double con = -(cur_stress.vol - p.residual_press);
    double scale = -con / (p.pref - p.residual_press);
    scale = pow(scale, p.pdc);
    G *= scale;
    K *= scale;
    for (int i = 1; i <= p.nsurf; ++i) {
      const double pm = modulus(i) * scale;
      const double sz = size(i) * con;
      t(kPSizes + i - 1) = sz;
      t(kPModuli + i - 1) = pm;
      set_center(i, zero6());
    }
  }

  __device__ void elastic_to_plastic() {
    if (p.load_stage != 1 || e2p == 1) return;
    e2p = 1;
    if (cur_stress.vol > 0. && p.phi > 0.) {
      const V6 d = cur_stress.dev;
      cur_stress.set(d, false);
    }
    param_scaling();
    if (cur_stress.dev_len() == 0.) return;
    while (yield(cur_stress, ++committed_active) > 0) {
      if (committed_active == p.nsurf) {
        deviator_scaling(cur_stress, p.nsurf);
        init_surface_update();
        return;
      }
    }
    committed_active--;
    init_surface_update();
  }

  __device__ void set_trial_stress(const T2& x) {
    V6 devia = x.dev;
    axpby(devia, 1.0, sub_rate.dev, 2. * G);
    trial.set_dev(devia, x.vol);
  }

  __device__ int set_sub_strain_rate() {
    const int n = p.nsurf;
    if (rate.is_zero()) return 0;
    double epm;
    if (active == 0) {
      epm = 2 * G;
    } else {
      const double pm = modulus(active);
      epm = 2 * G * pm / (2 * G + pm);
    }
    V6 inc = zero6();
    axpby(inc, 0.0, rate.dev, epm);
    T2 x;
    x.set(inc, false);
    const double single = size(n) / n;
    const double total = 3. * (sqrt(1. / 3.) * x.dev_len()) / sqrt(2.);
    int m = to_int(total / single + 1);
    if (m > n) m = n;
    inc = rate.t2;
    divide(inc, m);
    sub_rate.set(inc, false);
    return m;
  }

  __device__ V6 normal(const T2& x) const {
    V6 nrm = x.dev;
    const V6 cc = center(active);
    for (int k = 0; k < 6; ++k) nrm.v[k] -= cc.v[k];
    divide(nrm, sqrt(dd(nrm, nrm)));
    return nrm;
  }

  __device__ bool load_reversal() const {
    if (active == 0) return false;
    const V6 nrm = normal(cur_stress);
    V6 a = trial.dev;
    for (int k = 0; k < 6; ++k) a.v[k] -= cur_stress.dev.v[k];
    return dd(a, nrm) < 0;
  }

  __device__ T2 contact_stress() const {
    const V6 cc = center(active);
    V6 devia = trial.dev;
    for (int k = 0; k < 6; ++k) devia.v[k] -= cc.v[k];
    const double Ms = sqrt(3. / 2. * dd(devia, devia));
    const double f = size(active) / Ms;
    for (int k = 0; k < 6; ++k) devia.v[k] *= f;
    for (int k = 0; k < 6; ++k) devia.v[k] += cc.v[k];
    T2 x;
    x.set_dev(devia, trial.vol);
    return x;
  }

  __device__ double loading(const T2& cs, const V6& nrm, bool crossed) const {
    const double t1 = 2. * G, t2 = modulus(active);
    const double tt = t1 + t2;
    V6 tmp = trial.dev;
    for (int k = 0; k < 6; ++k) tmp.v[k] -= cs.dev.v[k];
    double lf = dd(nrm, tmp) / tt;
    if (crossed) {
      const double t3 = modulus(active - 1);
      lf *= (t3 - t2) / t3;
    }
    return lf;
  }

  __device__ bool crossing_next() const {
    if (active == p.nsurf) return false;
    return yield(trial, active + 1) > 0;
  }

  __device__ void stress_correction(bool crossed) {
    for (;;) {
      const T2 cs = contact_stress();
      const V6 nrm = normal(cs);
      const double lf = loading(cs, nrm, crossed);
      V6 devia = zero6();
      axpby(devia, 0.0, nrm, -2 * G * lf);
      for (int k = 0; k < 6; ++k) devia.v[k] += trial.dev.v[k];
      trial.set_dev(devia, trial.vol);
      deviator_scaling(trial, active);
      if (crossing_next()) {
        active++;
        crossed = true;
        continue;
      }
      return;
    }
  }

  __device__ void update_active_surface() {
    const int n = p.nsurf;
    if (active == n) return;
    V6 cen = center(active);
    const double sz = size(active);
    const V6 outc = center(active + 1);
    const double outsz = size(active + 1);
    V6 t1, t2, tmp;
    for (int k = 0; k < 6; ++k) {
      t1.v[k] = trial.dev.v[k] - cen.v[k];
      t2.v[k] = cen.v[k] - outc.v[k];
    }
    double A = dd(t1, t1);
    double B = 2. * dd(t1, t2);
    double Cc = dd(t2, t2) - 2. / 3. * outsz * outsz;
    double X = quad_root(A, B, Cc, ok);
    if (fabs(X - 1.) < kLow) X = 1.;
    if (X < 1.) {
      ok = false;
      return;
    }
    tmp = cen;
    axpby(tmp, 1.0, t1, X);
    const double f = 1.0 - sz / outsz;
    for (int k = 0; k < 6; ++k) tmp.v[k] *= f;
    t2 = cen;
    axpby(t2, 1.0, outc, -sz / outsz);
    for (int k = 0; k < 6; ++k) tmp.v[k] -= t2.v[k];
    T2 dir;
    dir.set(tmp, false);
    if (dir.dev_len() < kLow) return;
    tmp = dir.dev;
    A = dd(tmp, tmp);
    B = -2 * dd(t1, tmp);
    if (fabs(B) < kLow) B = 0.;
    Cc = dd(t1, t1) - 2. / 3. * sz * sz;
    if (fabs(Cc) < kLow || fabs(Cc) / dd(t1, t1) < kLow) return;
    if ((fabs(B) < 1.0e-10) && (fabs(Cc) < 1.0e-10)) return;
    if (B > 0. || Cc < 0.) {
      ok = false;
      return;
    }
    X = quad_root(A, B, Cc, ok);
    axpby(cen, 1.0, tmp, X);
    set_center(active, cen);
  }

  __device__ void update_inner_surface() {
    if (active <= 1) return;
    const V6 devia = cur_stress.dev;
    const V6 cen = center(active);
    const double sz = size(active);
    for (int i = 1; i < active; ++i) {
      V6 nc;
      const double f = size(i) / sz;
      for (int k = 0; k < 6; ++k) nc.v[k] = (cen.v[k] - devia.v[k]) * f + devia.v[k];
      set_center(i, nc);
    }
  }

  // the tangent's two parts: coeff and the deviator from the active surface's centre (getTangent
  // :860-886), zero when elastic
  __device__ void tangent_parts(double& coeff, V6& devia) const {
    coeff = 0.;
    devia = zero6();
    if (p.load_stage == 1 && active > 0) {
      const V6 cc = center(active);
      devia = trial.dev;
      for (int k = 0; k < 6; ++k) devia.v[k] -= cc.v[k];
      const double sz = size(active), pm = modulus(active);
      coeff = 6. * G * G / (2. * G + pm) / sz / sz;
    }
  }

  // entry (i, j) of the tangent from its parts (the operations of tangent below)
  __device__ static double tangent_entry(int load_stage, double coeff, const V6& devia, double G, double K, int i,
                                         int j) {
    double v = load_stage == 1 ? -coeff * devia.v[i] * devia.v[j] : 0.0;
    if (i == j) v += G;
    if (i < 3 && j < 3 && i == j) v += G;
    if (i < 3 && j < 3) v += (K - 2. * G / 3.);
    return v;
  }

  // the 6 x 6 tangent, row-major
  __device__ void tangent(double (&D)[36]) const {
    double coeff;
    V6 devia;
    tangent_parts(coeff, devia);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) D[6 * i + j] = tangent_entry(p.load_stage, coeff, devia, G, K, i, j);
  }

  // getStress (:455-509). With stop_at_correction (the compacted path's predictor) it returns
  // true, and stops, where the first stress correction would run; otherwise it sets the trial
  // stress and returns false.
  __device__ bool get_stress(const double (&eps)[6], bool stop_at_correction) {
    cur_stress = load_t2(kPStress);
    cur_strain = load_t2(kPStrain);
    committed_active = static_cast<int>(c(kPActive));
    active = committed_active;
    e2p = static_cast<int>(c(kPE2P));
    G = c(kPG);
    K = c(kPK);
    {
      V6 d, cs = cur_strain.engineering();
      for (int k = 0; k < 6; ++k) d.v[k] = eps[k] - cs.v[k];
      rate.set(d, true);
    }
    // committed surfaces into the trial buffer (centres of the nsurf surfaces in use)
    if (!fixed_surfaces())
      for (int k = kPSizes; k < kPCenters; ++k) t(k) = c(k);
    for (int k = kPCenters; k < kPCenters + 6 * ncopy; ++k) t(k) = c(k);
    if (p.load_stage == 1 && e2p == 0) elastic_to_plastic();
    if (p.load_stage != 1) {
      double De[36];
      tangent(De);
      V6 a = cur_stress.t2;
      const V6 r = rate.engineering();
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) a.v[j] += De[6 * j + i] * r.v[i];
      trial.set(a, false);
    } else {
      active = committed_active;
      sub_rate = rate;
      set_trial_stress(cur_stress);
      if (load_reversal()) {
        update_inner_surface();
        active = 0;
      }
      const int nsub = set_sub_strain_rate();
      for (int i = 0; i < nsub; ++i) {
        if (i == 0) set_trial_stress(cur_stress);
        else set_trial_stress(trial);
        if (active == 0 && !crossing_next()) continue;
        if (stop_at_correction) return true;
        if (active == 0) active++;
        stress_correction(false);
        update_active_surface();
      }
      double vol = K * (rate.vol * 3.);
      vol += cur_stress.vol;
      const V6 d = trial.dev;
      trial.set_dev(d, vol);
    }
    return false;
  }

  // finish (:512-539): the tangent, the outputs and the trial history.
  __device__ void finish(double (&sig)[6], double (&D)[36]) {
    tangent(D);
    for (int i = 0; i < 6; ++i) sig[i] = ok ? trial.t2.v[i] : NAN;
    V6 sc = cur_strain.t2;
    for (int k = 0; k < 6; ++k) sc.v[k] += rate.t2.v[k];
    T2 strain_commit;
    strain_commit.set(sc, false);
    store_t2(trial, kPStress);
    store_t2(strain_commit, kPStrain);
    t(kPE2P) = e2p;
    t(kPG) = G;
    t(kPK) = K;
    t(kPActive) = p.load_stage == 1 ? active : committed_active;
  }

  // the update at strain eps (engineering shear): stress, tangent, trial history written
  __device__ void update(const double (&eps)[6], double (&sig)[6], double (&D)[36]) {
    get_stress(eps, false);
    finish(sig, D);
  }

  // the same with the tangent as its parts (coeff, devia, G, K; entries by tangent_entry)
  __device__ void update_parts(const double (&eps)[6], double (&sig)[6], double& coeff, V6& devia) {
    get_stress(eps, false);
    tangent_parts(coeff, devia);
    double D[36];  // finish's tangent, unused (its stores fold away)
    finish(sig, D);
  }

  // The compacted path's predictor (pressure_independ_multi_yield.hpp:247-255): a point that
  // needs no stress correction is finished (false); otherwise true, and the corrector redoes
  // the whole update (the same function, so the bits are those of update).
  __device__ bool predict(const double (&eps)[6], double (&sig)[6], double (&D)[36]) {
    const bool plastic = get_stress(eps, true);
    if (!plastic) finish(sig, D);
    return plastic;
  }
};

}  // namespace pimy

// touched (may be null: every centre copied): the point's highest surface whose centre an update
// has written, read and raised here
__device__ inline void pimy_update(const PimyParams& p, const double* C, double* T, int64_t stride,
                                   const double (&eps)[6], double (&sig)[6], double (&D)[36],
                                   int32_t* touched = nullptr) {
  const int copy = touched ? *touched : -1;
  pimy::Point pt(p, C, T, stride, copy);
  pt.update(eps, sig, D);
  if (touched && pt.moved > copy) *touched = pt.moved;
}

// pimy_update with the tangent as its parts (PimyTangentParts below, load_stage p.load_stage)
struct PimyTangentParts;
__device__ inline void pimy_update_parts(const PimyParams& p, const double* C, double* T, int64_t stride,
                                         const double (&eps)[6], double (&sig)[6], PimyTangentParts& parts,
                                         int32_t* touched = nullptr);

// committed_tangent (:162-168): getTangent with the committed stress, active surface, surfaces
// and moduli (the surfaces are read from C; nothing is written).
__device__ inline void pimy_committed_tangent(const PimyParams& p, const double* C, int64_t stride,
                                              double (&D)[36]) {
  pimy::Point pt(p, C, const_cast<double*>(C), stride);
  pt.trial = pt.load_t2(kPStress);
  pt.active = static_cast<int>(pt.c(kPActive));
  pt.G = pt.c(kPG);
  pt.K = pt.c(kPK);
  pt.tangent(D);
}

// The committed tangent's parts (entries by pimy::Point::tangent_entry with p.load_stage, G, K):
// the damping force forms sigma_d entry by entry without holding the 6 x 6 matrix.
struct PimyTangentParts {
  double coeff, G, K;
  pimy::V6 devia;
  __device__ double entry(int load_stage, int i, int j) const {
    return pimy::Point::tangent_entry(load_stage, coeff, devia, G, K, i, j);
  }
};
__device__ inline void pimy_update_parts(const PimyParams& p, const double* C, double* T, int64_t stride,
                                         const double (&eps)[6], double (&sig)[6], PimyTangentParts& parts,
                                         int32_t* touched) {
  const int copy = touched ? *touched : -1;
  pimy::Point pt(p, C, T, stride, copy);
  pt.update_parts(eps, sig, parts.coeff, parts.devia);
  parts.G = pt.G;
  parts.K = pt.K;
  if (touched && pt.moved > copy) *touched = pt.moved;
}

__device__ inline PimyTangentParts pimy_committed_tangent_parts(const PimyParams& p, const double* C, int64_t stride) {
  pimy::Point pt(p, C, const_cast<double*>(C), stride);
  pt.trial = pt.load_t2(kPStress);
  pt.active = static_cast<int>(pt.c(kPActive));
  pt.G = pt.c(kPG);
  pt.K = pt.c(kPK);
  PimyTangentParts t;
  pt.tangent_parts(t.coeff, t.devia);
  t.G = pt.G;
  t.K = pt.K;
  return t;
}

__device__ inline bool pimy_predict(const PimyParams& p, const double* C, double* T, int64_t stride,
                                    const double (&eps)[6], double (&sig)[6], double (&D)[36],
                                    int32_t* touched = nullptr) {
  const int copy = touched ? *touched : -1;
  pimy::Point pt(p, C, T, stride, copy);
  const bool plastic = pt.predict(eps, sig, D);
  if (touched && pt.moved > copy) *touched = pt.moved;
  return plastic;
}

}  // namespace fx
