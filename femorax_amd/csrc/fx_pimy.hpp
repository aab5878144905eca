// PressureIndependMultiYield (the J2 multi-surface soil return map) as device code.
//
// The algorithm is the reference's (src/physics/materials/pressure_independ_multi_yield.hpp:
// update :228-240, getStress :455-509, finish :512-539, elast2Plast :542-571, yieldFunc
// :573-580, deviatorScaling :583-609, initSurfaceUpdate :611-631, paramScaling :633-658,
// setTrialStress :660-666, setSubStrainRate :668-692, getContactStress :703-712,
// isLoadReversal :714-722, getSurfaceNormal :724-730, getLoadingFunc :732-750,
// stressCorrection :753-774, updateActiveSurface :776-832, updateInnerSurface :834-849,
// isCrossingNextSurface :851-856, getTangent :860-886; tensor helpers t2_vector.hpp:32-220):
// its arithmetic and branch decisions are kept, because the stress path and the number of
// sub-increments depend on them.
//
// What is this build's own: one thread per material point; the point's history lives in
// global memory as structure-of-arrays, field k of point p at state[k * npts + p], so every
// field access of a wave is one coalesced load or store; the 6-vectors and the current, trial
// and rate tensors live in registers; the surfaces (sizes, moduli, centres) are read and
// written in place in the trial buffer. The file is compiled without multiply-add contraction
// so that the yield tests and sub-increment counts see the same roundings as the oracle.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace fx {

constexpr int kPimyNS = 20;                     // surfaces per point (automatic NS = 20)
constexpr int kPimyNState = 30 + 8 * kPimyNS;  // doubles per point
// state fields (the reference's layout, pressure_independ_multi_yield.hpp:74-85)
constexpr int kPStress = 0, kPStrain = 13, kPActive = 26, kPE2P = 27, kPG = 28, kPK = 29, kPSizes = 30,
              kPModuli = 30 + kPimyNS, kPCenters = 30 + 2 * kPimyNS;

struct PimyParams {
  int32_t load_stage, nsurf;
  double rho, G, K, phi, gamma_peak, pref, cohesion, pdc, residual_press;
  double b[3];
  int32_t mass_type, pad;
  double size[kPimyNS + 1], modulus[kPimyNS + 1];
  double alphaM, betaK, betaKinit, betaKcomm;  // the element's Rayleigh damping (batches)
};

}  // namespace fx
