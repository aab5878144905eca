// PressureIndependMultiYield parameter setup on the host: the backbone yield surfaces (sizes
// and plastic moduli) from (Gmax, cohesion, peak shear strain, NS), or from user (gamma,
// G/Gmax) pairs (pressure_independ_multi_yield.hpp: the constructor's checks :110-139 and
// setUpSurfaces :275-381). Runs once per parameter set.
#include <cmath>
#include <cstring>

#include "fx_internal.hpp"
#include "fx_pimy.hpp"

namespace fx {

int pimy_setup(const fx_pimy_material& m, PimyParams& p, std::string& err) {
  double rho = m.rho, G = m.G, K = m.K, cohesion = m.cohesion, gamma_peak = m.gamma_peak, phi = m.phi, pref = m.pref,
         pdc = m.pdc;
  int nsurf = m.nsurf;
  const double* gredu = m.gredu;
  if (G <= 0 || K <= 0) {
    err = "PressureIndependMultiYield: G and K must be positive";
    return FX_EINVAL;
  }
  if (phi < 0.) phi = 0.;
  if (phi == 0. && cohesion <= 0.) {
    err = "PressureIndependMultiYield: cohesion must be positive when the friction angle is 0";
    return FX_EINVAL;
  }
  if (cohesion <= 0) cohesion = 0.;
  if (gamma_peak <= 0 || pref <= 0) {
    err = "PressureIndependMultiYield: peak shear strain and reference pressure must be positive";
    return FX_EINVAL;
  }
  if (pdc < 0) pdc = 0.;
  if (pdc > 0 && phi == 0) pdc = 0.;
  if (nsurf <= 0) nsurf = 10;
  if (nsurf > 100) nsurf = 100;
  if (rho < 0) rho = 0.;
  if (nsurf > kPimyNS) {
    err = "PressureIndependMultiYield: at most 20 yield surfaces";
    return FX_EUNSUP;
  }
  if (gredu && nsurf < 2) {
    err = "PressureIndependMultiYield: user surfaces need at least 2 (gamma, G/Gmax) pairs";
    return FX_EINVAL;
  }
  std::memset(&p, 0, sizeof p);
  p.load_stage = 0;
  p.nsurf = nsurf;
  p.rho = rho;
  p.G = G;
  p.K = K;
  p.gamma_peak = gamma_peak;
  p.pref = -pref;  // compression negative
  p.pdc = pdc;
  for (int k = 0; k < 3; ++k) p.b[k] = m.b[k];
  p.mass_type = m.mass_type;
  const double pi = 3.14159265358979;
  double residual = 0., cone = 0., peak = 0., ref_strain = 0.;
  if (!gredu) {
    if (phi > 0) {
      const double sinPhi = std::sin(phi * pi / 180.);
      const double Mnys = 6. * sinPhi / (3. - sinPhi);
      residual = 3. * cohesion / (std::sqrt(2.) * Mnys);
      cone = -(p.pref - residual);
      peak = std::sqrt(2.) * cone * Mnys / 3.;
      ref_strain = (gamma_peak * peak) / (G * gamma_peak - peak);
    } else {
      peak = 2 * std::sqrt(2.) * cohesion / 3;
      ref_strain = (gamma_peak * peak) / (G * gamma_peak - peak);
      residual = 0.;
    }
    const double inc = peak / nsurf;
    for (int i = 1; i <= nsurf; ++i) {
      const double s1 = i * inc, s2 = s1 + inc;
      const double e1 = s1 * ref_strain / (G * ref_strain - s1);
      const double e2 = s2 * ref_strain / (G * ref_strain - s2);
      const double size = phi > 0. ? 3. * s1 / std::sqrt(2.) / cone : 3. * s1 / std::sqrt(2.);
      const double epm = 2. * (s2 - s1) / (e2 - e1);
      double pm = (2. * G - epm) <= 0 ? 1.0e30 : (2. * G * epm) / (2. * G - epm);
      if (pm < 0) pm = 0;
      if (pm > 1.0e30) pm = 1.0e30;
      if (i == nsurf) pm = 0;
      p.size[i] = size;
      p.modulus[i] = pm;
    }
  } else {
    const int last = 2 * (nsurf - 1);
    const double tmax = G * gredu[last] * gredu[last + 1];
    if (phi > 0) {
      const double Mnys = -(std::sqrt(3.) * tmax - 2. * cohesion) / p.pref;
      if (Mnys <= 0) {
        cohesion = std::sqrt(3.) / 2 * tmax;
        phi = 0.;
        cone = 1.;
        residual = 0.;
      } else {
        const double sinPhi = 3 * Mnys / (6 + Mnys);
        if (sinPhi < 0. || sinPhi > 1.) {
          err = "PressureIndependMultiYield: the user surfaces give an invalid friction angle";
          return FX_EINVAL;
        }
        residual = 2. * cohesion / Mnys;
        if (residual < 0.01 * p.pref) residual = 0.01 * p.pref;
        cone = -(p.pref - residual);
        phi = std::asin(sinPhi) * 180 / pi;
      }
    } else {
      cohesion = std::sqrt(3.) / 2 * tmax;
      cone = 1.;
      residual = 0.;
    }
    for (int i = 1; i < nsurf; ++i) {
      const int ii = 2 * (i - 1);
      const double e1 = gredu[ii], s1 = G * gredu[ii + 1] * e1;
      const double e2 = gredu[ii + 2], s2 = G * gredu[ii + 3] * e2;
      const double size = std::sqrt(3.) * s1 / cone;
      const double epm = 2. * (s2 - s1) / (e2 - e1);
      double pm = (2. * G - epm) <= 0 ? 1.0e30 : (2. * G * epm) / (2. * G - epm);
      if (pm <= 0) {
        err = "PressureIndependMultiYield: the user surfaces give a non-positive plastic modulus";
        return FX_EINVAL;
      }
      if (pm > 1.0e30) pm = 1.0e30;
      p.size[i] = size;
      p.modulus[i] = pm;
      if (i == nsurf - 1) {
        p.size[i + 1] = std::sqrt(3.) * s2 / cone;
        p.modulus[i + 1] = 0;
      }
    }
  }
  p.residual_press = residual;
  p.phi = phi;
  p.cohesion = cohesion;
  return FX_OK;
}

}  // namespace fx
