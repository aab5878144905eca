/*
 * fxo stdBrick + ElasticIsotropic + element operator — TEST INFRASTRUCTURE (the oracle).
 *
 * Restates src/physics/elements/std_brick.hpp (shp3d :45-152, point_geometry :203-211,
 * point_strain :276-301, point_node_force :305-327, body force :331-333, block tangent
 * :336-356, mass apply/entry :361-383, compute_with :493-543, mass :582-605, lumped_mass
 * :609-623), src/physics/materials/elastic_isotropic.hpp (:40-94) and
 * src/physics/operators/element_operator.hpp (Factors :45-56, effective_tangent :60-75,
 * apply :82-130, diagonal :138-209).
 *
 * The arithmetic keeps the reference's operation order (sums left to right in the same term
 * order, the same products), because the reference's golden files (OpenSees output) are
 * matched bit for bit by that order on a compiler that does not contract multiply-adds; this
 * file is built with -ffp-contract=off.
 */
#include <math.h>
#include <string.h>

#include "fxo.h"

/* Natural-coordinate signs of the 8 nodes (OpenSees Brick order). */
static const int SX[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
static const int SY[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
static const int SZ[8] = {-1, -1, -1, -1, +1, +1, +1, +1};

/* One factor (1 + s*xi) as shp3d forms it: ap = 1 + xi, am = 1 - xi. */
static inline double fac(int s, double xi) { return s > 0 ? 1.0 + xi : 1.0 - xi; }

void fxo_brick_shape(const double xl[3][8], int gp, double shp[4][8], double* dvol) {
  const double r3 = 1.0 / sqrt(3.0);
  const double sg[2] = {-r3, r3};
  const double ss[3] = {sg[(gp >> 2) & 1], sg[(gp >> 1) & 1], sg[gp & 1]};
  /* natural derivatives and values: dN/dxi = sx * (0.125*fy)*fz, etc. (shp3d's c1, c2, c3) */
  for (int k = 0; k < 8; ++k) {
    const double fx = fac(SX[k], ss[0]), fy = fac(SY[k], ss[1]), fz = fac(SZ[k], ss[2]);
    const double cyz = 0.125 * fy * fz; /* c2 */
    const double cxz = 0.125 * fx * fz; /* c3 */
    const double cxy = 0.125 * fx * fy; /* c1 */
    shp[0][k] = SX[k] > 0 ? cyz : -cyz;
    shp[1][k] = SY[k] > 0 ? cxz : -cxz;
    shp[2][k] = SZ[k] > 0 ? cxy : -cxy;
    shp[3][k] = cxy * fz;
  }
  /* Jacobian from the edge differences, in shp3d's pairing and term order */
  double xs[3][3], ad[3][3];
  for (int j = 0; j < 3; ++j) {
    const double* x = xl[j];
    xs[j][0] = (x[1] - x[0]) * shp[0][1] + (x[2] - x[3]) * shp[0][2] + (x[5] - x[4]) * shp[0][5] +
               (x[6] - x[7]) * shp[0][6];
    xs[j][1] = (x[2] - x[1]) * shp[1][2] + (x[3] - x[0]) * shp[1][3] + (x[6] - x[5]) * shp[1][6] +
               (x[7] - x[4]) * shp[1][7];
    xs[j][2] = (x[4] - x[0]) * shp[2][4] + (x[5] - x[1]) * shp[2][5] + (x[6] - x[2]) * shp[2][6] +
               (x[7] - x[3]) * shp[2][7];
  }
  /* adjugate by cyclic cofactors, determinant along the first row */
  ad[0][0] = xs[1][1] * xs[2][2] - xs[1][2] * xs[2][1];
  ad[0][1] = xs[2][1] * xs[0][2] - xs[2][2] * xs[0][1];
  ad[0][2] = xs[0][1] * xs[1][2] - xs[0][2] * xs[1][1];
  ad[1][0] = xs[1][2] * xs[2][0] - xs[1][0] * xs[2][2];
  ad[1][1] = xs[2][2] * xs[0][0] - xs[2][0] * xs[0][2];
  ad[1][2] = xs[0][2] * xs[1][0] - xs[0][0] * xs[1][2];
  ad[2][0] = xs[1][0] * xs[2][1] - xs[1][1] * xs[2][0];
  ad[2][1] = xs[2][0] * xs[0][1] - xs[2][1] * xs[0][0];
  ad[2][2] = xs[0][0] * xs[1][1] - xs[0][1] * xs[1][0];
  const double det = xs[0][0] * ad[0][0] + xs[0][1] * ad[1][0] + xs[0][2] * ad[2][0];
  const double rdet = 1.0 / det;
  double inv[3][3]; /* inv[i][j] = ad[i][j] / det */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) inv[i][j] = ad[i][j] * rdet;
  for (int k = 0; k < 8; ++k) {
    const double a = shp[0][k], b = shp[1][k], c = shp[2][k];
    shp[0][k] = a * inv[0][0] + b * inv[1][0] + c * inv[2][0];
    shp[1][k] = a * inv[0][1] + b * inv[1][1] + c * inv[2][1];
    shp[2][k] = a * inv[0][2] + b * inv[1][2] + c * inv[2][2];
  }
  *dvol = 1.0 * det; /* weight 1 */
}

static void local_coords(const double xe[8][3], double xl[3][8]) {
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) xl[d][i] = xe[i][d];
}

/* strain 11 22 33 12 23 31 (engineering shear) from nodal values (std_brick.hpp:276-301) */
static void point_strain(const double shp[4][8], const double* ul, double eps[6]) {
  for (int c = 0; c < 6; ++c) eps[c] = 0.0;
  for (int j = 0; j < 8; ++j) {
    const double dx = shp[0][j], dy = shp[1][j], dz = shp[2][j];
    const double u0 = ul[3 * j], u1 = ul[3 * j + 1], u2 = ul[3 * j + 2];
    eps[0] += dx * u0;
    eps[1] += dy * u1;
    eps[2] += dz * u2;
    eps[3] += dy * u0 + dx * u1;
    eps[4] += dz * u1 + dy * u2;
    eps[5] += dz * u0 + dx * u2;
  }
}

/* B_j^T (sigma * dvol) (std_brick.hpp:305-327) */
static void point_node_force(const double shp[4][8], double dvol, const double sig[6], int j,
                             double out[3]) {
  double s[6];
  for (int c = 0; c < 6; ++c) s[c] = sig[c] * dvol;
  const double dx = shp[0][j], dy = shp[1][j], dz = shp[2][j];
  out[0] = dx * s[0] + dy * s[3] + dz * s[5];
  out[1] = dy * s[1] + dx * s[3] + dz * s[4];
  out[2] = dz * s[2] + dy * s[4] + dx * s[5];
}

/* B_J^T (D * dvol) B_K, OpenSees' two loops (std_brick.hpp:336-356, computeB :698-710) */
static void fill_B(const double shp[4][8], int node, double B[6][3]) {
  memset(B, 0, sizeof(double) * 18);
  B[0][0] = shp[0][node];
  B[1][1] = shp[1][node];
  B[2][2] = shp[2][node];
  B[3][0] = shp[1][node];
  B[3][1] = shp[0][node];
  B[4][1] = shp[2][node];
  B[4][2] = shp[1][node];
  B[5][0] = shp[2][node];
  B[5][2] = shp[0][node];
}
static void point_block_tangent(const double shp[4][8], double dvol, const double* D, int J, int K,
                                double out[3][3]) {
  double BJ[6][3], BK[6][3], BD[3][6];
  fill_B(shp, J, BJ);
  fill_B(shp, K, BK);
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 6; ++q) {
      double sum = 0.0;
      for (int r = 0; r < 6; ++r) sum += BJ[r][p] * (D[6 * r + q] * dvol);
      BD[p][q] = sum;
    }
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      double sum = 0.0;
      for (int r = 0; r < 6; ++r) sum += BD[p][r] * BK[r][q];
      out[p][q] = sum;
    }
}

void fxo_elastic_tangent(double E, double nu, double D[6][6]) {
  double mu2 = E / (1.0 + nu);
  const double lam = nu * mu2 / (1.0 - 2.0 * nu);
  const double mu = 0.50 * mu2;
  mu2 += lam;
  memset(D, 0, sizeof(double) * 36);
  D[0][0] = D[1][1] = D[2][2] = mu2;
  D[0][1] = D[1][0] = D[0][2] = D[2][0] = D[1][2] = D[2][1] = lam;
  D[3][3] = D[4][4] = D[5][5] = mu;
}

/* ElasticIsotropic::update: stress from the explicit formulas (elastic_isotropic.hpp:40-63) */
static void elastic_stress(double E, double nu, const double eps[6], double sig[6]) {
  double mu2 = E / (1.0 + nu);
  const double lam = nu * mu2 / (1.0 - 2.0 * nu);
  const double mu = 0.50 * mu2;
  mu2 += lam;
  sig[0] = mu2 * eps[0] + lam * (eps[1] + eps[2]);
  sig[1] = mu2 * eps[1] + lam * (eps[0] + eps[2]);
  sig[2] = mu2 * eps[2] + lam * (eps[0] + eps[1]);
  sig[3] = mu * eps[3];
  sig[4] = mu * eps[4];
  sig[5] = mu * eps[5];
}

void fxo_brick_force_elastic(const fxo_elastic* p, const double xe[8][3], const double* ue,
                             const double* ve, double* fe, double* Ke) {
  if (p->element == 1) {
    fxo_ssp_force(p, xe, 0, ue, ve, fe, Ke);
    return;
  }
  double xl[3][8];
  local_coords(xe, xl);
  double resid[24], fdamp[24];
  for (int a = 0; a < 24; ++a) resid[a] = fdamp[a] = 0.0;
  if (Ke) memset(Ke, 0, sizeof(double) * 576);
  double D[6][6], D0[6][6];
  fxo_elastic_tangent(p->E, p->nu, D);
  memcpy(D0, D, sizeof D);
  const fxo_rayleigh* r = &p->damp;
  const int stiff_damp = r->betaK != 0.0 || r->betaKinit != 0.0 || r->betaKcomm != 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol;
    fxo_brick_shape(xl, gp, shp, &dvol);
    double eps[6], sig[6], dsig[6] = {0, 0, 0, 0, 0, 0}, vgp[3] = {0, 0, 0};
    point_strain(shp, ue, eps);
    elastic_stress(p->E, p->nu, eps, sig);
    if (stiff_damp) {
      double rate[6];
      point_strain(shp, ve, rate);
      for (int a = 0; a < 6; ++a) {
        double sum = 0.0;
        for (int c = 0; c < 6; ++c) {
          double Dac = r->betaK * D[a][c];
          if (r->betaKinit != 0.0) Dac += r->betaKinit * D0[a][c];
          if (r->betaKcomm != 0.0) Dac += r->betaKcomm * D[a][c]; /* committed = D (elastic) */
          sum += Dac * rate[c];
        }
        dsig[a] = sum;
      }
    }
    if (r->alphaM != 0.0 && p->mass_type == 0)
      for (int j = 0; j < 8; ++j)
        for (int q = 0; q < 3; ++q) vgp[q] += shp[3][j] * ve[3 * j + q];
    for (int j = 0; j < 8; ++j) {
      double rj[3], dj[3] = {0, 0, 0}, body[3];
      point_node_force(shp, dvol, sig, j, rj);
      for (int q = 0; q < 3; ++q) body[q] = dvol * p->b[q] * shp[3][j];
      if (stiff_damp) point_node_force(shp, dvol, dsig, j, dj);
      if (r->alphaM != 0.0) {
        const double temp = shp[3][j] * dvol * p->rho;
        for (int q = 0; q < 3; ++q) {
          const double vq = p->mass_type == 0 ? vgp[q] : ve[3 * j + q];
          dj[q] += r->alphaM * temp * vq;
        }
      }
      for (int q = 0; q < 3; ++q) {
        resid[3 * j + q] += rj[q];
        resid[3 * j + q] -= body[q];
        fdamp[3 * j + q] += dj[q];
      }
    }
    if (Ke)
      for (int J = 0; J < 8; ++J)
        for (int K = 0; K < 8; ++K) {
          double blk[3][3];
          point_block_tangent(shp, dvol, &D[0][0], J, K, blk);
          for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) Ke[(3 * J + a) * 24 + 3 * K + c] += blk[a][c];
        }
  }
  for (int a = 0; a < 24; ++a) fe[a] = resid[a] + fdamp[a];
}

void fxo_brick_mass(const fxo_elastic* p, const double xe[8][3], double* Me) {
  double xl[3][8];
  local_coords(xe, xl);
  memset(Me, 0, sizeof(double) * 576);
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol;
    fxo_brick_shape(xl, gp, shp, &dvol);
    for (int j = 0; j < 8; ++j) {
      double temp = shp[3][j] * dvol;
      temp *= p->rho;
      for (int k = 0; k < 8; ++k) {
        double mjk;
        if (p->mass_type == 0) mjk = temp * shp[3][k];
        else if (j == k) mjk = temp;
        else continue;
        for (int q = 0; q < 3; ++q) Me[(3 * j + q) * 24 + 3 * k + q] += mjk;
      }
    }
  }
}

void fxo_brick_lumped_mass(double rho, const double xe[8][3], double m[8]) {
  double xl[3][8];
  local_coords(xe, xl);
  for (int j = 0; j < 8; ++j) m[j] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol;
    fxo_brick_shape(xl, gp, shp, &dvol);
    for (int j = 0; j < 8; ++j) {
      double temp = shp[3][j] * dvol;
      temp *= rho;
      m[j] += temp;
    }
  }
}

/* Factors of element_operator.hpp:45-56 (stdBrick: initial stiffness is not its tangent) */
static void factors(const fxo_rayleigh* d, const fxo_coeffs* k, double* cK, double* c0, double* cc,
                    double* cM) {
  *cK = k->initial ? k->damping * d->betaK : k->stiffness + k->damping * d->betaK;
  *c0 = k->initial ? k->stiffness + k->damping * d->betaKinit : k->damping * d->betaKinit;
  *cc = k->damping * d->betaKcomm;
  *cM = k->mass + k->damping * d->alphaM;
}

/* effective_tangent (element_operator.hpp:60-75); for elastic D0 = Dc = D */
static void effective(double cK, double c0, double cc, const double* D, double* Deff) {
  for (int a = 0; a < 36; ++a) {
    double v = cK * D[a];
    if (c0 != 0.0) v += c0 * D[a];
    if (cc != 0.0) v += cc * D[a];
    Deff[a] = v;
  }
}

/* point_mass_apply (std_brick.hpp:361-374) */
static void point_mass_apply(const fxo_elastic* p, const double shp[4][8], double dvol,
                             const double* v, int j, double out[3]) {
  const double temp = shp[3][j] * dvol * p->rho;
  for (int q = 0; q < 3; ++q) {
    double vq;
    if (p->mass_type == 0) {
      vq = 0.0;
      for (int k = 0; k < 8; ++k) vq += shp[3][k] * v[3 * k + q];
    } else {
      vq = v[3 * j + q];
    }
    out[q] = temp * vq;
  }
}

void fxo_brick_operator_apply(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3],
                              const double* D, const double* ve, double* ye) {
  if (p->element == 1) {
    fxo_ssp_operator_apply(p, k, xe, ve, ye);
    return;
  }
  double cK, c0, cc, cM;
  factors(&p->damp, k, &cK, &c0, &cc, &cM);
  double xl[3][8], Deff[36];
  local_coords(xe, xl);
  effective(cK, c0, cc, D, Deff);
  for (int a = 0; a < 24; ++a) ye[a] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol, eps[6], sig[6];
    fxo_brick_shape(xl, gp, shp, &dvol);
    point_strain(shp, ve, eps);
    for (int a = 0; a < 6; ++a) {
      double sum = 0.0;
      for (int c = 0; c < 6; ++c) sum += Deff[6 * a + c] * eps[c];
      sig[a] = sum;
    }
    for (int j = 0; j < 8; ++j) {
      double fj[3];
      point_node_force(shp, dvol, sig, j, fj);
      for (int q = 0; q < 3; ++q) ye[3 * j + q] += fj[q];
      if (cM != 0.0) {
        double mj[3];
        point_mass_apply(p, shp, dvol, ve, j, mj);
        for (int q = 0; q < 3; ++q) ye[3 * j + q] += cM * mj[q];
      }
    }
  }
}

static double point_mass_entry(const fxo_elastic* p, const double shp[4][8], double dvol, int j, int k) {
  const double temp = shp[3][j] * dvol * p->rho;
  if (p->mass_type == 0) return temp * shp[3][k];
  return j == k ? temp : 0.0;
}

void fxo_brick_operator_diagonal(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3],
                                 const double* D, const int32_t* eq, double* de) {
  if (p->element == 1) {
    fxo_ssp_operator_diagonal(p, k, xe, eq, de);
    return;
  }
  double cK, c0, cc, cM;
  factors(&p->damp, k, &cK, &c0, &cc, &cM);
  double xl[3][8], Deff[36];
  local_coords(xe, xl);
  effective(cK, c0, cc, D, Deff);
  for (int a = 0; a < 24; ++a) de[a] = 0.0;
  int first[24], merged = 0;
  for (int a = 0; a < 24; ++a) {
    first[a] = a;
    if (eq[a] < 0) continue;
    for (int b = 0; b < a; ++b)
      if (eq[b] == eq[a]) {
        first[a] = first[b];
        merged = 1;
        break;
      }
  }
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol;
    fxo_brick_shape(xl, gp, shp, &dvol);
    for (int j = 0; j < 8; ++j) {
      double blk[3][3];
      point_block_tangent(shp, dvol, Deff, j, j, blk);
      const double m = cM != 0.0 ? cM * point_mass_entry(p, shp, dvol, j, j) : 0.0;
      for (int q = 0; q < 3; ++q) de[3 * j + q] += blk[q][q] + m;
    }
    if (!merged) continue;
    for (int J = 0; J < 8; ++J)
      for (int K = 0; K < 8; ++K) {
        int any = 0;
        for (int a = 0; a < 3 && !any; ++a)
          for (int c = 0; c < 3; ++c) {
            const int ia = 3 * J + a, ic = 3 * K + c;
            if (ia != ic && eq[ia] >= 0 && eq[ia] == eq[ic]) any = 1;
          }
        if (!any) continue;
        double blk[3][3];
        point_block_tangent(shp, dvol, Deff, J, K, blk);
        const double m = (cM != 0.0 && J != K) ? cM * point_mass_entry(p, shp, dvol, J, K) : 0.0;
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) {
            const int ia = 3 * J + a, ic = 3 * K + c;
            if (ia == ic || eq[ia] < 0 || eq[ia] != eq[ic]) continue;
            de[first[ia]] += blk[a][c] + (a == c ? m : 0.0);
          }
      }
  }
}

/* ---- stdBrick with PressureIndependMultiYield ------------------------------------------------ */
/* the elastic initial tangent of the material's reference moduli (initial_tangent) */
void fxo_pimy_initial_tangent(const fxo_pimy_params* m, double D[36]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double v = 0.;
      if (i == j) v += m->G;
      if (i < 3 && j < 3 && i == j) v += m->G;
      if (i < 3 && j < 3) v += (m->K - 2. * m->G / 3.);
      D[6 * i + j] = v;
    }
}
static void pimy_initial_D(const fxo_pimy_elem* p, double* D) { fxo_pimy_initial_tangent(&p->mat, D); }

/* Brick::compute_with (std_brick.hpp:493-543) with the material update per Gauss point
 * (point_update_shape :409-443, point_node_contribution :448-464): per point and node, resid +=
 * B_j^T sigma dvol; resid -= b N_j dvol; fdamp += B_j^T sigma_d dvol + alphaM N_j dvol rho v,
 * sigma_d = (betaK D + betaKinit D_0 + betaKcomm D_c) B v with D the point's updated tangent,
 * D_0 the initial and D_c the committed tangent; fe = resid + fdamp. C[gp], T[gp]: the point's
 * committed and trial history; Dgp: the points' tangents (8 x 36) or NULL; ve NULL: no damping
 * velocity (fe = resid). */
void fxo_brick_force_pimy(const fxo_pimy_elem* p, const double xe[8][3], const double* ue, const double* ve,
                          const double* const* C, double* const* T, double* fe, double* Dgp) {
  if (p->element == 1) {
    fxo_ssp_force_pimy(p, xe, ue, ve, C[0], T[0], fe, Dgp);
    return;
  }
  const fxo_rayleigh* dp = &p->damp;
  const int stiff = dp->betaK != 0.0 || dp->betaKinit != 0.0 || dp->betaKcomm != 0.0;
  const int damped = ve && (stiff || dp->alphaM != 0.0);
  double xl[3][8], D0[36];
  local_coords(xe, xl);
  if (damped && dp->betaKinit != 0.0) pimy_initial_D(p, D0);
  double resid[24], fdamp[24];
  for (int a = 0; a < 24; ++a) resid[a] = fdamp[a] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol, eps[6], sig[6], D[36];
    fxo_brick_shape(xl, gp, shp, &dvol);
    point_strain(shp, ue, eps);
    fxo_pimy_update(&p->mat, C[gp], T[gp], eps, sig, D);
    if (Dgp) memcpy(Dgp + 36 * gp, D, sizeof D);
    double dsig[6] = {0, 0, 0, 0, 0, 0}, vgp[3] = {0, 0, 0};
    if (damped && stiff) {
      double Dc[36], rate[6];
      if (dp->betaKcomm != 0.0) fxo_pimy_committed_tangent(&p->mat, C[gp], Dc);
      point_strain(shp, ve, rate);
      for (int a = 0; a < 6; ++a) {
        double sum = 0.0;
        for (int c = 0; c < 6; ++c) {
          double Dac = dp->betaK * D[6 * a + c];
          if (dp->betaKinit != 0.0) Dac += dp->betaKinit * D0[6 * a + c];
          if (dp->betaKcomm != 0.0) Dac += dp->betaKcomm * Dc[6 * a + c];
          sum += Dac * rate[c];
        }
        dsig[a] = sum;
      }
    }
    if (damped && dp->alphaM != 0.0 && p->mass_type == 0)
      for (int j = 0; j < 8; ++j)
        for (int q = 0; q < 3; ++q) vgp[q] += shp[3][j] * ve[3 * j + q];
    for (int j = 0; j < 8; ++j) {
      double rj[3], body[3], dj[3] = {0, 0, 0};
      point_node_force(shp, dvol, sig, j, rj);
      for (int q = 0; q < 3; ++q) body[q] = dvol * p->b[q] * shp[3][j];
      if (damped) {
        if (stiff) point_node_force(shp, dvol, dsig, j, dj);
        if (dp->alphaM != 0.0) {
          const double temp = shp[3][j] * dvol * p->mat.rho;
          for (int q = 0; q < 3; ++q) dj[q] += dp->alphaM * temp * (p->mass_type == 0 ? vgp[q] : ve[3 * j + q]);
        }
      }
      for (int q = 0; q < 3; ++q) {
        resid[3 * j + q] += rj[q];
        resid[3 * j + q] -= body[q];
        fdamp[3 * j + q] += dj[q];
      }
    }
  }
  for (int a = 0; a < 24; ++a) fe[a] = damped ? resid[a] + fdamp[a] : resid[a];
}

/* The operator factors (element_operator.hpp:45-56; the PIMY initial stiffness is not its
 * tangent): cK on the stored tangent, c0 on the initial one, cc on the committed one, cM on the
 * mass. */
static void pimy_factors(const fxo_pimy_elem* p, const fxo_coeffs* k, double* cK, double* c0, double* cc,
                         double* cM) {
  const fxo_rayleigh* d = &p->damp;
  *cK = k->initial ? k->damping * d->betaK : k->stiffness + k->damping * d->betaK;
  *c0 = k->initial ? k->stiffness + k->damping * d->betaKinit : k->damping * d->betaKinit;
  *cc = k->damping * d->betaKcomm;
  *cM = k->mass + k->damping * d->alphaM;
}

/* D_eff at point gp (effective_tangent, element_operator.hpp:58-75) */
static void pimy_effective(const fxo_pimy_elem* p, double cK, double c0, double cc, const double* Dgp,
                           const double* D0, const double* const* C, int gp, double* Deff) {
  double Dc[36];
  if (cc != 0.0) fxo_pimy_committed_tangent(&p->mat, C[gp], Dc);
  for (int a = 0; a < 36; ++a) {
    double v = cK * Dgp[36 * gp + a];
    if (c0 != 0.0) v += c0 * D0[a];
    if (cc != 0.0) v += cc * Dc[a];
    Deff[a] = v;
  }
}

/* element_operator_apply with stored tangents (Source::Stored, element_operator.hpp:82-130) */
void fxo_brick_operator_apply_gp(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3],
                                 const double* Dgp, const double* const* C, const double* ve, double* ye) {
  if (p->element == 1) {
    fxo_ssp_operator_apply_pimy(p, k, xe, Dgp, C ? C[0] : NULL, ve, ye);
    return;
  }
  double cK, c0, cc, cM;
  pimy_factors(p, k, &cK, &c0, &cc, &cM);
  double xl[3][8], D0[36];
  local_coords(xe, xl);
  pimy_initial_D(p, D0);
  fxo_elastic me;
  memset(&me, 0, sizeof me);
  me.rho = p->mat.rho;
  me.mass_type = p->mass_type;
  for (int a = 0; a < 24; ++a) ye[a] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol, eps[6], sig[6], Deff[36];
    fxo_brick_shape(xl, gp, shp, &dvol);
    pimy_effective(p, cK, c0, cc, Dgp, D0, C, gp, Deff);
    point_strain(shp, ve, eps);
    for (int a = 0; a < 6; ++a) {
      double sum = 0.0;
      for (int c = 0; c < 6; ++c) sum += Deff[6 * a + c] * eps[c];
      sig[a] = sum;
    }
    for (int j = 0; j < 8; ++j) {
      double fj[3];
      point_node_force(shp, dvol, sig, j, fj);
      for (int q = 0; q < 3; ++q) ye[3 * j + q] += fj[q];
      if (cM != 0.0) {
        double mj[3];
        point_mass_apply(&me, shp, dvol, ve, j, mj);
        for (int q = 0; q < 3; ++q) ye[3 * j + q] += cM * mj[q];
      }
    }
  }
}

/* element_operator_diagonal with stored tangents, merged DOFs' cross terms on the first */
void fxo_brick_operator_diagonal_gp(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3],
                                    const double* Dgp, const double* const* C, const int32_t* eq, double* de) {
  if (p->element == 1) {
    fxo_ssp_operator_diagonal_pimy(p, k, xe, Dgp, C ? C[0] : NULL, eq, de);
    return;
  }
  double cK, c0, cc, cM;
  pimy_factors(p, k, &cK, &c0, &cc, &cM);
  double xl[3][8], D0[36];
  local_coords(xe, xl);
  pimy_initial_D(p, D0);
  fxo_elastic me;
  memset(&me, 0, sizeof me);
  me.rho = p->mat.rho;
  me.mass_type = p->mass_type;
  for (int a = 0; a < 24; ++a) de[a] = 0.0;
  int first[24], merged = 0;
  for (int a = 0; a < 24; ++a) {
    first[a] = a;
    if (eq[a] < 0) continue;
    for (int b = 0; b < a; ++b)
      if (eq[b] == eq[a]) {
        first[a] = first[b];
        merged = 1;
        break;
      }
  }
  for (int gp = 0; gp < 8; ++gp) {
    double shp[4][8], dvol, Deff[36];
    fxo_brick_shape(xl, gp, shp, &dvol);
    pimy_effective(p, cK, c0, cc, Dgp, D0, C, gp, Deff);
    for (int j = 0; j < 8; ++j) {
      double blk[3][3];
      point_block_tangent(shp, dvol, Deff, j, j, blk);
      const double m = cM != 0.0 ? cM * point_mass_entry(&me, shp, dvol, j, j) : 0.0;
      for (int q = 0; q < 3; ++q) de[3 * j + q] += blk[q][q] + m;
    }
    if (!merged) continue;
    for (int J = 0; J < 8; ++J)
      for (int K = 0; K < 8; ++K) {
        int any = 0;
        for (int a = 0; a < 3 && !any; ++a)
          for (int c = 0; c < 3; ++c) {
            const int ia = 3 * J + a, ic = 3 * K + c;
            if (ia != ic && eq[ia] >= 0 && eq[ia] == eq[ic]) any = 1;
          }
        if (!any) continue;
        double blk[3][3];
        point_block_tangent(shp, dvol, Deff, J, K, blk);
        const double m = (cM != 0.0 && J != K) ? cM * point_mass_entry(&me, shp, dvol, J, K) : 0.0;
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) {
            const int ia = 3 * J + a, ic = 3 * K + c;
            if (ia == ic || eq[ia] < 0 || eq[ia] != eq[ic]) continue;
            de[first[ia]] += blk[a][c] + (a == c ? m : 0.0);
          }
      }
  }
}
