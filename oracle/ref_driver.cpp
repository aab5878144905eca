// oracle/_ref driver — TEST INFRASTRUCTURE.
//
// Compiles the reference's own header-only physics (src/physics/elements/std_brick.hpp,
// materials/elastic_isotropic.hpp, materials/pressure_independ_multi_yield.hpp,
// operators/element_operator.hpp) straight from /root/reference/src and exports a few C entry
// points, so the tests can check the C restatement (oracle/fxo_*.c) against the reference's
// own code on inputs the golden files do not cover. Built by oracle/Makefile (target `ref`)
// into oracle/_ref/libfxref.so; nothing of the reference is copied. The physics headers need
// neither Kokkos nor any other library (they are plain C++20).
#include <cstring>

#include "physics/elements/std_brick.hpp"
#include "physics/materials/elastic_isotropic.hpp"
#include "physics/materials/pressure_independ_multi_yield.hpp"
#include "physics/operators/element_operator.hpp"

using namespace femorax;
using Brick = StdBrick<ElasticIsotropic>;

namespace {
Brick::Params brick_params(const double* mat /* E nu rho */, const double* b, int mass_type) {
  Brick::Params p{};
  p.mat = ElasticIsotropic::Params{mat[0], mat[1], mat[2]};
  p.b[0] = b[0];
  p.b[1] = b[1];
  p.b[2] = b[2];
  p.massType = mass_type;
  return p;
}
Rayleigh rayleigh(const double* r) {
  Rayleigh d{};
  if (r) {
    d.alphaM = r[0];
    d.betaK = r[1];
    d.betaKinit = r[2];
    d.betaKcomm = r[3];
  }
  return d;
}
}  // namespace

extern "C" {

// Brick::compute<true> (std_brick.hpp:547-553): fe[24], Ke[576].
void fxref_brick_force(const double* mat, const double* b, int mass_type, const double* ray,
                       const double* xe, const double* u, const double* v, double* fe, double* Ke) {
  const Brick::Params p = brick_params(mat, b, mass_type);
  double x[8][3];
  std::memcpy(x, xe, sizeof x);
  double K[24][24];
  Brick::compute<true>(p, rayleigh(ray), x, u, v, {}, {}, fe, K);
  if (Ke) std::memcpy(Ke, K, sizeof K);
}

void fxref_brick_mass(const double* mat, int mass_type, const double* xe, double* Me) {
  const double b[3] = {0, 0, 0};
  const Brick::Params p = brick_params(mat, b, mass_type);
  double x[8][3];
  std::memcpy(x, xe, sizeof x);
  Brick::mass(p, x, Me);
}

// element_operator_apply / _diagonal with the recomputed geometry and the constant elastic D.
void fxref_operator(const double* mat, int mass_type, const double* ray, const double* coeffs /* k m c */,
                    int initial, const double* xe, const double* v, const int* eq, int diagonal, double* out) {
  const double b[3] = {0, 0, 0};
  const Brick::Params p = brick_params(mat, b, mass_type);
  const Rayleigh d = rayleigh(ray);
  OperatorCoefficients k;
  k.stiffness = coeffs[0];
  k.damping = coeffs[1];
  k.mass = coeffs[2];
  k.initial = initial != 0;
  double x[8][3], xl[3][8];
  std::memcpy(x, xe, sizeof x);
  Brick::local_coords(x, xl);
  const auto geo = Brick::recomputed_geometry(xl, p);
  auto tangent_at = [&](int, const Brick::PointShape&, double (*D)[6]) { ElasticIsotropic::initial_tangent(p.mat, D); };
  auto committed_at = [&](int, double (*D)[6]) { ElasticIsotropic::initial_tangent(p.mat, D); };
  if (diagonal) element_operator_diagonal<Brick>(p, d, k, geo, tangent_at, committed_at, eq, out);
  else element_operator_apply<Brick>(p, d, k, geo, tangent_at, committed_at, v, out);
}

// PressureIndependMultiYield<20>: setup, then one update from a committed buffer.
int fxref_pimy_setup(const double* in /* rho G K c gamma phi pref pdc */, int nsurf, const double* pairs,
                     double* params_out /* sizeof(Params)/8 doubles */) {
  using M = PressureIndependMultiYield<20>;
  M::Input i{};
  i.rho = in[0];
  i.refShearModul = in[1];
  i.refBulkModul = in[2];
  i.cohesi = in[3];
  i.peakShearStra = in[4];
  i.frictionAng = in[5];
  i.refPress = in[6];
  i.pressDependCoe = in[7];
  i.numberOfYieldSurf = nsurf;
  i.gredu = pairs;
  M::Params p{};
  const int rc = M::setup(i, p);
  std::memcpy(params_out, &p, sizeof p);
  return rc;
}
int fxref_pimy_params_bytes() { return static_cast<int>(sizeof(PressureIndependMultiYield<20>::Params)); }
int fxref_pimy_nstate() { return PressureIndependMultiYield<20>::NSTATE; }

void fxref_pimy_initial_state(const double* params, double* state) {
  using M = PressureIndependMultiYield<20>;
  M::Params p;
  std::memcpy(&p, params, sizeof p);
  M::initial_state(p, {state, 1});
}

void fxref_pimy_update(const double* params, int load_stage, const double* committed, double* trial,
                       const double* eps, double* sig, double* D) {
  using M = PressureIndependMultiYield<20>;
  M::Params p;
  std::memcpy(&p, params, sizeof p);
  p.loadStage = load_stage;
  double DD[6][6];
  M::update(p, {committed, 1}, {trial, 1}, eps, sig, DD);
  std::memcpy(D, DD, sizeof DD);
}

// StdBrick<PressureIndependMultiYield<20>>::compute<false> (std_brick.hpp:493-543): the element
// force with the return map at each point and Rayleigh damping at v; histories contiguous per
// point (NSTATE each), committed C, trial T written.
void fxref_pimy_brick_force(const double* params, int load_stage, const double* b, int mass_type, const double* ray,
                            const double* xe, const double* u, const double* v, const double* C, double* T,
                            double* fe) {
  using M = PressureIndependMultiYield<20>;
  using PB = StdBrick<M>;
  PB::Params p{};
  std::memcpy(&p.mat, params, sizeof p.mat);
  p.mat.loadStage = load_stage;
  p.b[0] = b[0];
  p.b[1] = b[1];
  p.b[2] = b[2];
  p.massType = mass_type;
  double x[8][3];
  std::memcpy(x, xe, sizeof x);
  const PointArray<const double> committed{C, 1, M::NSTATE};
  const PointArray<double> trial{T, 1, M::NSTATE};
  PB::compute<false>(p, rayleigh(ray), x, u, v, committed, trial, fe, nullptr);
}

// PressureIndependMultiYield<20>::committed_tangent (:162-168)
void fxref_pimy_committed_tangent(const double* params, int load_stage, const double* committed, double* D) {
  using M = PressureIndependMultiYield<20>;
  M::Params p;
  std::memcpy(&p, params, sizeof p);
  p.loadStage = load_stage;
  double DD[6][6];
  M::committed_tangent(p, {committed, 1}, DD);
  std::memcpy(D, DD, sizeof DD);
}

// element_operator_apply / _diagonal of StdBrick<PressureIndependMultiYield<20>> with the points'
// tangents Dgp[8][36] (the stored source) and the committed tangent from C (contiguous points).
void fxref_pimy_operator(const double* params, int load_stage, int mass_type, const double* ray,
                         const double* coeffs /* k c m */, int initial, const double* xe, const double* Dgp,
                         const double* C, const double* v, const int* eq, int diagonal, double* out) {
  using M = PressureIndependMultiYield<20>;
  using PB = StdBrick<M>;
  PB::Params p{};
  std::memcpy(&p.mat, params, sizeof p.mat);
  p.mat.loadStage = load_stage;
  p.massType = mass_type;
  const Rayleigh d = rayleigh(ray);
  OperatorCoefficients k;
  k.stiffness = coeffs[0];
  k.damping = coeffs[1];
  k.mass = coeffs[2];
  k.initial = initial != 0;
  double x[8][3], xl[3][8];
  std::memcpy(x, xe, sizeof x);
  PB::local_coords(x, xl);
  const auto geo = PB::recomputed_geometry(xl, p);
  auto tangent_at = [&](int gp, const PB::PointShape&, double (*D)[6]) {
    std::memcpy(D, Dgp + 36 * gp, 36 * sizeof(double));
  };
  auto committed_at = [&](int gp, double (*D)[6]) {
    M::committed_tangent(p.mat, {C + static_cast<std::ptrdiff_t>(gp) * M::NSTATE, 1}, D);
  };
  if (diagonal) element_operator_diagonal<PB>(p, d, k, geo, tangent_at, committed_at, eq, out);
  else element_operator_apply<PB>(p, d, k, geo, tangent_at, committed_at, v, out);
}

// StdBrick<PressureIndependMultiYield<20>>::critical_dt at the committed state C (contiguous points):
// the stability tangent of each point (stability_moduli, :174-203).
double fxref_pimy_critical_dt(const double* params, int load_stage, const double* ray, int scheme, const double* xe,
                              const double* C) {
  using M = PressureIndependMultiYield<20>;
  using PB = StdBrick<M>;
  PB::Params p{};
  std::memcpy(&p.mat, params, sizeof p.mat);
  p.mat.loadStage = load_stage;
  double x[8][3];
  std::memcpy(x, xe, sizeof x);
  const PointArray<const double> committed{C, 1, M::NSTATE};
  return PB::critical_dt(p, rayleigh(ray), scheme == 0 ? ExplicitScheme::CentralDifference
                                                      : ExplicitScheme::ExplicitDifference, x, committed);
}

// StdBrick::critical_dt (std_brick.hpp:629-694) with ElasticIsotropic (no state):
// scheme 0 CentralDifference, 1 ExplicitDifference. Returns omega^2 in *omega_sq.
double fxref_brick_critical_dt(const double* mat, const double* ray, int scheme, const double* xe,
                               double* omega_sq) {
  const double b[3] = {0, 0, 0};
  const Brick::Params p = brick_params(mat, b, 0);
  double x[8][3];
  std::memcpy(x, xe, sizeof x);
  const PointArray<const double> none{};
  *omega_sq = Brick::max_omega_sq(p, x, none);
  return Brick::critical_dt(p, rayleigh(ray), scheme == 0 ? ExplicitScheme::CentralDifference
                                                         : ExplicitScheme::ExplicitDifference, x, none);
}

}  // extern "C"
