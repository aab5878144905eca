// Host-side SSPbrick constants, computed once per batch (the reference's geometry cache and
// element constants, element_kernels.hpp:47, built by SSPbrick::point_store_geometry and
// element_constants, ssp_brick.hpp:890-914).
//
// SSPbrick (stabilized single-point hexahedron, McGann, Arduino and Mackenzie-Helnwein) is one
// mean-strain point plus a stabilisation stiffness Kstab = Mben^T interior Mben:
//   - the mean-strain rows dNmod of node i, the volume mVol and the Jacobian determinant terms
//     J[0..19] come from the seven moment vectors e_k = (x, y, z) . {xi, et, ze, hut, hus, hst,
//     hstu} of the node coordinates (ssp_brick.hpp:141-221);
//   - Mben's rows are the hourglass vectors gamma_g = (I - dNmod x^T) h_g (:223-249);
//   - interior = FCF - FeCFhg^T FeCFe^-1 FeCFhg (:265-796) is written here in closed form: the
//     reference's integrals are moments of the Jacobian determinant polynomial over the natural
//     cube, one per monomial, and the hourglass modes hst = xi eta, hut = eta zeta, hus = xi zeta,
//     hstu = xi eta zeta have monomial gradients, so every block is an isotropic contraction
//     (C1 = lambda + 2 mu, C2 = lambda, C3 = mu) of those moments with Jinv. Blocks that involve
//     hstu carry the shear modulus only and FeCFhg's hstu columns are zero, as in the reference.
//     FeCFe^-1 FeCFhg comes from a Cholesky solve (FeCFe is symmetric positive definite for a
//     positive-definite elastic tangent; a failed factorisation leaves interior = FCF, as the
//     reference's zeroed inverse does).
// The node volume weights (mass and body force) are the integrals of N_i (nodal_volumes,
// :800-811): the stdBrick lumped weights.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fx_internal.hpp"

namespace fx {

namespace {

constexpr int kSX[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
constexpr int kSY[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
constexpr int kSZ[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// 0.125 x the sign pattern of vector v at node i: xi et ze hst hut hus hstu
double pattern(int v, int i) {
  int s = 0;
  switch (v) {
    case 0: s = kSX[i]; break;
    case 1: s = kSY[i]; break;
    case 2: s = kSZ[i]; break;
    case 3: s = kSX[i] * kSY[i]; break;
    case 4: s = kSY[i] * kSZ[i]; break;
    case 5: s = kSX[i] * kSZ[i]; break;
    default: s = kSX[i] * kSY[i] * kSZ[i]; break;
  }
  return 0.125 * s;
}

struct V3 {
  double x, y, z;
  double operator[](int d) const { return d == 0 ? x : d == 1 ? y : z; }
};
V3 cross(const V3& a, const V3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double triple(const V3& a, const V3& b, const V3& c) { return dot(a, cross(b, c)); }

// The moment over the natural cube of xi^a eta^b zeta^c times the Jacobian determinant polynomial,
// in the reference's J combinations (ssp_brick.hpp:270-289).
double moment(const double (&J)[20], int a, int b, int c) {
  switch (100 * a + 10 * b + c) {
    case 200: return 8.0 * (J[0] / 3.0 + J[7] / 5.0 + J[8] / 9.0 + J[9] / 9.0);
    case 20: return 8.0 * (J[0] / 3.0 + J[8] / 5.0 + J[7] / 9.0 + J[9] / 9.0);
    case 2: return 8.0 * (J[0] / 3.0 + J[9] / 5.0 + J[7] / 9.0 + J[8] / 9.0);
    case 11: return 8.0 * (J[4] / 9.0 + J[17] / 27.0);
    case 101: return 8.0 * (J[5] / 9.0 + J[18] / 27.0);
    case 110: return 8.0 * (J[6] / 9.0 + J[19] / 27.0);
    case 102: return 8.0 * (J[1] / 9.0 + J[12] / 15.0 + J[15] / 27.0);
    case 120: return 8.0 * (J[1] / 9.0 + J[15] / 15.0 + J[12] / 27.0);
    case 12: return 8.0 * (J[2] / 9.0 + J[11] / 15.0 + J[14] / 27.0);
    case 210: return 8.0 * (J[2] / 9.0 + J[14] / 15.0 + J[11] / 27.0);
    case 21: return 8.0 * (J[3] / 9.0 + J[10] / 15.0 + J[13] / 27.0);
    case 201: return 8.0 * (J[3] / 9.0 + J[13] / 15.0 + J[10] / 27.0);
    case 220: return 8.0 * (J[0] / 9.0 + J[7] / 15.0 + J[8] / 15.0 + J[9] / 27.0);
    case 22: return 8.0 * (J[0] / 9.0 + J[8] / 15.0 + J[9] / 15.0 + J[7] / 27.0);
    case 202: return 8.0 * (J[0] / 9.0 + J[7] / 15.0 + J[9] / 15.0 + J[8] / 27.0);
    case 211: return 8.0 * (J[4] / 27.0 + 64.0 * J[17] / 45.0);
    case 121: return 8.0 * (J[5] / 27.0 + 64.0 * J[18] / 45.0);
    case 112: return 8.0 * (J[6] / 27.0 + 64.0 * J[19] / 45.0);
    case 111: return 8.0 * J[16] / 27.0;
    default: return 0.0;
  }
}

// exponents of hourglass mode m (hst, hut, hus, hstu) and of its derivative along natural axis a
constexpr int kMode[4][3] = {{1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
bool derivative(int m, int a, int (&ex)[3]) {
  if (!kMode[m][a]) return false;
  for (int k = 0; k < 3; ++k) ex[k] = kMode[m][k] - (k == a);
  return true;
}

// isotropic pairing of two gradients g, h for DOF directions d, e
double iso(const double* g, const double* h, double C1, double C2, double C3, int d, int e) {
  if (d != e) return C2 * g[d] * h[e] + C3 * g[e] * h[d];
  double s = 0.0;
  for (int f = 0; f < 3; ++f)
    if (f != d) s += g[f] * h[f];
  return C1 * g[d] * h[d] + C3 * s;
}

}  // namespace

void ssp_constants(const double (&xe)[8][3], double lam, double mu, double* c, int64_t stride) {
  // moment vectors e1..e7: xi et ze hut hus hst hstu
  constexpr int kVec[7] = {0, 1, 2, 4, 5, 3, 6};
  V3 e[8];
  for (int k = 0; k < 7; ++k) {
    double m[3] = {0, 0, 0};
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < 3; ++d) m[d] += xe[i][d] * pattern(kVec[k], i);
    e[k + 1] = {m[0], m[1], m[2]};
  }
  double J[20] = {};
  J[0] = triple(e[1], e[2], e[3]);
  J[4] = triple(e[7], e[2], e[3]) + triple(e[4], e[5], e[2]) + triple(e[4], e[3], e[6]);
  J[5] = triple(e[1], e[7], e[3]) + triple(e[4], e[5], e[1]) + triple(e[3], e[5], e[6]);
  J[6] = triple(e[1], e[2], e[7]) + triple(e[4], e[1], e[6]) + triple(e[2], e[5], e[6]);
  J[7] = -triple(e[1], e[5], e[6]);
  J[8] = -triple(e[4], e[2], e[6]);
  J[9] = -triple(e[4], e[5], e[3]);
  J[10] = triple(e[2], e[4], e[7]);
  J[11] = -triple(e[3], e[4], e[7]);
  J[12] = triple(e[3], e[5], e[7]);
  J[13] = -triple(e[1], e[5], e[7]);
  J[14] = triple(e[1], e[6], e[7]);
  J[15] = -triple(e[2], e[6], e[7]);
  J[16] = 2.0 * triple(e[4], e[5], e[6]);
  J[17] = triple(e[7], e[5], e[6]);
  J[18] = triple(e[4], e[7], e[6]);
  J[19] = triple(e[4], e[5], e[7]);
  const double vol = 8.0 * (J[0] + (J[7] + J[8] + J[9]) / 3.0);
  // mean-strain rows
  const V3 lin[3] = {cross(e[2], e[3]), cross(e[3], e[1]), cross(e[1], e[2])};
  const V3 quad[3] = {cross(e[6], e[5]), cross(e[4], e[6]), cross(e[5], e[4])};
  const V3 hq[3] = {cross(e[5], e[1]), cross(e[6], e[2]), cross(e[1], e[6])};   // hst hut hus
  const V3 hr[3] = {cross(e[2], e[4]), cross(e[3], e[5]), cross(e[4], e[3])};
  double dN[8][3];
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) {
      double l = 0.0, q = 0.0;
      for (int a = 0; a < 3; ++a) {
        l += lin[a][d] * pattern(a, i);
        q += quad[a][d] * pattern(a, i);
      }
      q += (hq[0][d] + hr[0][d]) * pattern(3, i) + (hq[1][d] + hr[1][d]) * pattern(4, i) +
           (hq[2][d] + hr[2][d]) * pattern(5, i);
      dN[i][d] = (8.0 * l + (8.0 / 3.0) * q) / vol;
    }
  // Jinv: the inverse of Jmat[d][a] = e_{a+1}[d] (adjugate over determinant)
  double Jinv[3][3] = {};
  {
    const double A[3][3] = {{e[1].x, e[2].x, e[3].x}, {e[1].y, e[2].y, e[3].y}, {e[1].z, e[2].z, e[3].z}};
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (det != 0.0) {
      Jinv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
      Jinv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
      Jinv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
      Jinv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
      Jinv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
      Jinv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
      Jinv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
      Jinv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
      Jinv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    }
  }
  // hourglass vectors gamma_g(i) = h_g(i) - sum_k dN[i][k] (x_k . h_g)
  double gam[4][8];
  for (int m = 0; m < 4; ++m) {
    double xh[3] = {0, 0, 0};
    for (int j = 0; j < 8; ++j)
      for (int k = 0; k < 3; ++k) xh[k] += xe[j][k] * pattern(3 + m, j);
    for (int i = 0; i < 8; ++i)
      gam[m][i] = pattern(3 + m, i) - (dN[i][0] * xh[0] + dN[i][1] * xh[1] + dN[i][2] * xh[2]);
  }
  const double C1 = lam + 2.0 * mu, C2 = lam, C3 = mu;
  // FCF (12 x 12, blocks of hourglass modes)
  double F[12][12];
  for (int g = 0; g < 4; ++g)
    for (int h = 0; h < 4; ++h) {
      double P[3][3] = {};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          int ea[3], eb[3];
          if (!derivative(g, a, ea) || !derivative(h, b, eb)) continue;
          const double I = moment(J, ea[0] + eb[0], ea[1] + eb[1], ea[2] + eb[2]);
          for (int X = 0; X < 3; ++X)
            for (int Y = 0; Y < 3; ++Y) P[X][Y] += I * Jinv[a][X] * Jinv[b][Y];
        }
      const bool shear = g == 3 || h == 3;
      for (int d = 0; d < 3; ++d)
        for (int f = 0; f < 3; ++f) {
          double v;
          if (d == f) {
            double s = 0.0;
            for (int o = 0; o < 3; ++o)
              if (o != d) s += P[o][o];
            v = (shear ? 0.0 : C1 * P[d][d]) + C3 * s;
          } else {
            v = shear ? C3 * P[f][d] : C2 * P[d][f] + C3 * P[f][d];
          }
          F[3 * g + d][3 * h + f] = v;
        }
    }
  // FeCFe (9 x 9) and FeCFhg (9 x 12, hstu columns zero)
  double E[9][9], H[9][12] = {};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      int ex[3] = {0, 0, 0};
      ++ex[a];
      ++ex[b];
      const double I = moment(J, ex[0], ex[1], ex[2]);
      for (int d = 0; d < 3; ++d)
        for (int f = 0; f < 3; ++f) E[3 * a + d][3 * b + f] = iso(Jinv[a], Jinv[b], C1, C2, C3, d, f) * I;
    }
  for (int a = 0; a < 3; ++a)
    for (int g = 0; g < 3; ++g)
      for (int b = 0; b < 3; ++b) {
        int eb[3];
        if (!derivative(g, b, eb)) continue;
        ++eb[a];
        const double I = moment(J, eb[0], eb[1], eb[2]);
        for (int d = 0; d < 3; ++d)
          for (int f = 0; f < 3; ++f) H[3 * a + d][3 * g + f] += iso(Jinv[a], Jinv[b], C1, C2, C3, d, f) * I;
      }
  // Cholesky E = L L^T, then X = E^-1 H and interior = F - H^T X
  double L[9][9] = {};
  bool ok = true;
  for (int i = 0; i < 9 && ok; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = E[i][j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(s > 0.0)) {
          ok = false;
          break;
        }
        L[i][i] = std::sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  double X[9][12] = {};
  if (ok)
    for (int col = 0; col < 12; ++col) {
      double y[9];
      for (int i = 0; i < 9; ++i) {
        double s = H[i][col];
        for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
        y[i] = s / L[i][i];
      }
      for (int i = 8; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 9; ++k) s -= L[k][i] * X[k][col];
        X[i][col] = s / L[i][i];
      }
    }
  // store: dNmod (24), mVol, node weights (8), gamma (32), interior's upper triangle (78)
  for (int k = 0; k < 24; ++k) c[k * stride] = dN[k / 3][k % 3];
  c[24 * stride] = vol;
  double w[8];
  brick_node_weights(xe, w);
  for (int j = 0; j < 8; ++j) c[(25 + j) * stride] = w[j];
  for (int m = 0; m < 4; ++m)
    for (int i = 0; i < 8; ++i) c[(33 + 8 * m + i) * stride] = gam[m][i];
  int k = 65;
  for (int r = 0; r < 12; ++r)
    for (int q = r; q < 12; ++q) {
      double s = 0.0;
      for (int t = 0; t < 9; ++t) s += H[t][r] * X[t][q];
      c[(k++) * stride] = F[r][q] - s;
    }
}

namespace {
double interior(const double* c, int r, int q) {
  if (r > q) std::swap(r, q);
  return c[65 + 12 * r - r * (r - 1) / 2 + (q - r)];
}
}  // namespace

void ssp_stiffness(const double* c, double lam, double mu, double (&K)[24][24]) {
  // Kstab(a, b) = sum_gh gamma_g(i) interior(3g + d, 3h + e) gamma_h(k), a = 3i + d, b = 3k + e
  for (int a = 0; a < 24; ++a)
    for (int b = 0; b < 24; ++b) {
      const int i = a / 3, d = a % 3, k = b / 3, e = b % 3;
      double s = 0.0;
      for (int g = 0; g < 4; ++g)
        for (int h = 0; h < 4; ++h) s += c[33 + 8 * g + i] * interior(c, 3 * g + d, 3 * h + e) * c[33 + 8 * h + k];
      K[a][b] = s;
    }
  // + mVol Bnot_i^T D Bnot_k
  const double V = c[24];
  for (int i = 0; i < 8; ++i)
    for (int k = 0; k < 8; ++k) {
      const double* gi = c + 3 * i;
      const double* gk = c + 3 * k;
      const double gg = gi[0] * gk[0] + gi[1] * gk[1] + gi[2] * gk[2];
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e)
          K[3 * i + d][3 * k + e] += V * (lam * gi[d] * gk[e] + mu * gi[e] * gk[d] + (d == e ? mu * gg : 0.0));
    }
}

double ssp_max_omega_sq(const double (&xe)[8][3], double lam, double mu, double rho) {
  double c[kSspConsts], K[24][24], s[8];
  ssp_constants(xe, lam, mu, c, 1);
  for (int j = 0; j < 8; ++j) {
    const double m = rho * c[25 + j];
    if (!(m > 0.0)) return 0.0;  // massless: no limit
    s[j] = 1.0 / std::sqrt(m);
  }
  ssp_stiffness(c, lam, mu, K);
  for (int a = 0; a < 24; ++a)
    for (int b = 0; b < 24; ++b) K[a][b] *= s[a / 3] * s[b / 3];
  return max_eigenvalue_symmetric(K);
}

void ssp_box_constants(double hx, double hy, double hz, double lam, double mu, double* out) {
  const double box[8][3] = {{0, 0, 0}, {hx, 0, 0}, {hx, hy, 0}, {0, hy, 0},
                            {0, 0, hz}, {hx, 0, hz}, {hx, hy, hz}, {0, hy, hz}};
  double c[kSspConsts], K[24][24];
  ssp_constants(box, lam, mu, c, 1);
  // the stabilisation in the modal basis of the trilinear field: Mben v = (H v)_m / 8 at the
  // hourglass modes m = 3 (xi eta), 6 (eta zeta), 5 (xi zeta), 7 (xi eta zeta), so the modal
  // operator is interior / 64
  int k = 0;
  for (int r = 0; r < 12; ++r)
    for (int q = r; q < 12; ++q) out[k++] = interior(c, r, q) / 64.0;
  // diag(K_e) by lattice corner q = i + 2j + 4k (reference node order 0 1 3 2 4 5 7 6)
  ssp_stiffness(c, lam, mu, K);
  constexpr int kNodeOfCorner[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  for (int d = 0; d < 3; ++d)
    for (int q = 0; q < 8; ++q) out[78 + 8 * d + q] = K[3 * kNodeOfCorner[q] + d][3 * kNodeOfCorner[q] + d];
}

}  // namespace fx
