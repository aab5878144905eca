// Host-side stdBrick quantities computed once per batch: nodal lumped-mass weights (the body
// load and the lumped mass share them), the element stiffness for the stable time step, and the
// largest eigenvalue of a small symmetric matrix.
//
// Reference: lumped mass rho * sum_gp N_j dvol (std_brick.hpp:609-623), body force
// b * sum_gp N_j dvol (std_brick.hpp:519-542), the stable step from the largest omega^2 of
// M_l^-1/2 K M_l^-1/2 (std_brick.hpp:658-694) by cyclic Jacobi rotations (sym_eigen.hpp:11-49;
// Golub & Van Loan, Matrix Computations, Algorithm 8.4.1), and stable_dt
// (explicit_stability.hpp:15-32). 2 x 2 x 2 Gauss points at +-1/sqrt(3), weight 1.
#include <cmath>

#include "fx_internal.hpp"

namespace fx {

namespace {

constexpr double kSX[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
constexpr double kSY[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
constexpr double kSZ[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// N, dN/dx and det J at Gauss point gp = 4i + 2j + k.
void shape_at(const double (&xe)[8][3], int gp, double (&N)[8], double (&dx)[8][3], double& det) {
  const double g = 1.0 / std::sqrt(3.0);
  const double xi = ((gp >> 2) & 1 ? g : -g), eta = ((gp >> 1) & 1 ? g : -g), zeta = (gp & 1 ? g : -g);
  double dn[8][3];
  for (int a = 0; a < 8; ++a) {
    const double fx = 1 + kSX[a] * xi, fy = 1 + kSY[a] * eta, fz = 1 + kSZ[a] * zeta;
    N[a] = 0.125 * fx * fy * fz;
    dn[a][0] = 0.125 * kSX[a] * fy * fz;
    dn[a][1] = 0.125 * fx * kSY[a] * fz;
    dn[a][2] = 0.125 * fx * fy * kSZ[a];
  }
  double J[3][3] = {};
  for (int a = 0; a < 8; ++a)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) J[r][c] += dn[a][c] * xe[a][r];  // dx_r / dxi_c
  det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double inv[3][3];  // dxi_c / dx_r
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  for (int a = 0; a < 8; ++a)
    for (int r = 0; r < 3; ++r) dx[a][r] = dn[a][0] * inv[0][r] + dn[a][1] * inv[1][r] + dn[a][2] * inv[2][r];
}

}  // namespace

void brick_node_weights(const double (&xe)[8][3], double (&w)[8]) {
  for (int j = 0; j < 8; ++j) w[j] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double N[8], dx[8][3], det;
    shape_at(xe, gp, N, dx, det);
    for (int j = 0; j < 8; ++j) w[j] += N[j] * det;
  }
}

void brick_stiffness(const double (&xe)[8][3], double lam, double mu, double (&K)[24][24]) {
  const double L[8] = {lam, lam, lam, lam, lam, lam, lam, lam}, M[8] = {mu, mu, mu, mu, mu, mu, mu, mu};
  brick_stiffness_points(xe, L, M, K);
}

// K with an isotropic elastic tangent per Gauss point (lam[gp], mu[gp])
void brick_stiffness_points(const double (&xe)[8][3], const double (&lamp)[8], const double (&mup)[8],
                            double (&K)[24][24]) {
  for (auto& row : K)
    for (double& v : row) v = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    const double lam = lamp[gp], mu = mup[gp];
    double N[8], g[8][3], det;
    shape_at(xe, gp, N, g, det);
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) {
        const double gg = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
        for (int c = 0; c < 3; ++c)
          for (int d = 0; d < 3; ++d)
            K[3 * a + c][3 * b + d] += det * (lam * g[a][c] * g[b][d] + mu * g[a][d] * g[b][c] + (c == d ? mu * gg : 0.0));
      }
  }
}

double max_eigenvalue_symmetric(double (&A)[24][24]) {
  constexpr int n = 24;
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
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        A[p][q] = A[q][p] = 0.0;
      }
  }
  double lmax = A[0][0];
  for (int i = 1; i < n; ++i) lmax = std::max(lmax, A[i][i]);
  return lmax;
}

// the consistent mass of unit density, per node pair: sum_gp N_J N_K dvol (std_brick.hpp:361-383)
void brick_mass_unit(const double (&xe)[8][3], double (&M)[8][8]) {
  for (auto& row : M)
    for (double& v : row) v = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double N[8], dx[8][3], det;
    shape_at(xe, gp, N, dx, det);
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b) M[a][b] += N[a] * N[b] * det;
  }
}

double brick_max_omega_sq(const double (&xe)[8][3], double lam, double mu, double rho) {
  const double L[8] = {lam, lam, lam, lam, lam, lam, lam, lam}, M[8] = {mu, mu, mu, mu, mu, mu, mu, mu};
  return brick_max_omega_sq_points(xe, L, M, rho);
}

double brick_max_omega_sq_points(const double (&xe)[8][3], const double (&lamp)[8], const double (&mup)[8],
                                 double rho) {
  double w[8], K[24][24], s[8];
  brick_node_weights(xe, w);
  for (int j = 0; j < 8; ++j) {
    const double m = rho * w[j];
    if (!(m > 0.0)) return 0.0;  // massless: no limit
    s[j] = 1.0 / std::sqrt(m);
  }
  brick_stiffness_points(xe, lamp, mup, K);
  for (int a = 0; a < 24; ++a)
    for (int c = 0; c < 24; ++c) K[a][c] *= s[a / 3] * s[c / 3];
  return max_eigenvalue_symmetric(K);
}

double stable_dt(int scheme, const DevParam& p, double omega_sq) {
  if (!(omega_sq > 0.0)) return INFINITY;
  if (scheme == 0) return 2.0 / std::sqrt(omega_sq);
  const double beta = p.betaK + p.betaKinit + p.betaKcomm;
  const double c = p.alphaM + beta * omega_sq;
  return 4.0 / (std::sqrt(c * c + 4.0 * omega_sq) + c);
}

}  // namespace fx
