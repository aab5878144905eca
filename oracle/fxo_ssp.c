/*
 * fxo SSPbrick + ElasticIsotropic — TEST INFRASTRUCTURE (the oracle).
 *
 * Restates src/physics/elements/ssp_brick.hpp: the shape and hourglass vectors (ShapeVectors
 * :64-90), GetStab's geometry (geometry :141-249: the Jacobian determinant terms J[0..19], mVol,
 * the mean-strain rows dNmod, Jinv and the hourglass vectors gst gut gus gstu), GetStab's
 * stabilisation matrix (interior_matrix :265-796: FCF - KuT FeCFe^-1 FeCFhg), the exact nodal
 * volumes (nodal_volumes :800-811) and OpenSees' polyJac formula (poly_bracket :252-259),
 * element_constants (:898-914), element_term_reduce / _node / _entry (:964-1005), the point
 * functions (:1009-1069), compute_with (:1142-1193), mass (:1236-1249), max_omega_sq_with
 * (:1309-1340), and the element operator and diagonal with an element term
 * (src/physics/operators/element_operator.hpp:45-56, 82-130, 138-209;
 * src/physics/common/element_term.hpp:19-37).
 *
 * The 12 x 12 stabilisation matrix is written here in a compact form rather than as the
 * reference's explicit terms: every integral the reference's FCF / FeCFe / FeCFhg use is the
 * integral over the natural cube of a monomial xi^i eta^j zeta^k times the Jacobian
 * determinant polynomial, and the reference gives each of them as one of nineteen J
 * combinations (J0789 = the integral of xi^2, J619 = of xi eta, J174 = of xi^2 eta zeta, ...).
 * The hourglass modes hst = xi eta, hut = eta zeta, hus = xi zeta, hstu = xi eta zeta have
 * monomial gradients, so each block is an isotropic contraction of those integrals with
 * Jinv; the blocks that involve hstu keep only the shear modulus (C3), as the reference's do,
 * and FeCFhg's hstu columns are zero, as the reference's are. Pinned by the reference's
 * SSPbrick goldens (OpenSees output: force, tangent, damping force, mass on four geometries
 * and two materials) to 1e-12 of the largest entry.
 */
#include <math.h>
#include <string.h>

#include "fxo.h"

static const int SX[8] = {-1, +1, +1, -1, -1, +1, +1, -1};
static const int SY[8] = {-1, -1, +1, +1, -1, -1, +1, +1};
static const int SZ[8] = {-1, -1, -1, -1, +1, +1, +1, +1};

/* The vectors of SSPbrick::setDomain, node i: xi et ze hst hut hus hstu (0.125 times the signs). */
enum { V_XI, V_ET, V_ZE, V_HST, V_HUT, V_HUS, V_HSTU };
static double shape_vector(int v, int i) {
  const int s[7] = {SX[i], SY[i], SZ[i], SX[i] * SY[i], SY[i] * SZ[i], SX[i] * SZ[i], SX[i] * SY[i] * SZ[i]};
  return 0.125 * s[v];
}

static double dot8v(const double x[8], int v) {
  double r = 0.0;
  for (int i = 0; i < 8; ++i) r += x[i] * shape_vector(v, i);
  return r;
}

static void cross3(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
/* a . (b x c) */
static double triple3(const double a[3], const double b[3], const double c[3]) {
  double t[3];
  cross3(b, c, t);
  return a[0] * t[0] + a[1] * t[1] + a[2] * t[2];
}

/* Gauss-Jordan inverse with partial pivoting; 0 when singular (the reference's invert_small then
 * leaves zeros). */
static int invert_n(int n, const double* A, double* inv) {
  double M[9][18];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j) M[i][j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
    if (M[piv][c] == 0.0) return 0;
    if (piv != c)
      for (int j = 0; j < 2 * n; ++j) {
        const double t = M[c][j];
        M[c][j] = M[piv][j];
        M[piv][j] = t;
      }
    const double d = M[c][c];
    for (int j = 0; j < 2 * n; ++j) M[c][j] /= d;
    for (int r = 0; r < n; ++r) {
      if (r == c || M[r][c] == 0.0) continue;
      const double f = M[r][c];
      for (int j = 0; j < 2 * n; ++j) M[r][j] -= f * M[c][j];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) inv[i * n + j] = M[i][j + n];
  return 1;
}

typedef struct ssp_geom {
  double J[20];
  double mVol;
  double dN[8][3]; /* dNmod: node i's mean-strain row (bx, by, bz) */
  double Jinv[3][3];
  double gam[4][8]; /* gst gut gus gstu */
} ssp_geom;

/* GetStab's geometry (ssp_brick.hpp:141-249). */
static void ssp_geometry(const double xl[3][8], ssp_geom* g) {
  /* e_k = (x, y, z) . vector: e1 xi, e2 et, e3 ze, e4 hut, e5 hus, e6 hst, e7 hstu */
  static const int vec[7] = {V_XI, V_ET, V_ZE, V_HUT, V_HUS, V_HST, V_HSTU};
  double e[8][3];
  for (int k = 0; k < 7; ++k)
    for (int d = 0; d < 3; ++d) e[k + 1][d] = dot8v(xl[d], vec[k]);
  double* J = g->J;
  const double* e1 = e[1];
  const double* e2 = e[2];
  const double* e3 = e[3];
  const double* e4 = e[4];
  const double* e5 = e[5];
  const double* e6 = e[6];
  const double* e7 = e[7];
  J[0] = triple3(e1, e2, e3);
  J[1] = J[2] = J[3] = 0.0;
  J[4] = triple3(e7, e2, e3) + triple3(e4, e5, e2) + triple3(e4, e3, e6);
  J[5] = triple3(e1, e7, e3) + triple3(e4, e5, e1) + triple3(e3, e5, e6);
  J[6] = triple3(e1, e2, e7) + triple3(e4, e1, e6) + triple3(e2, e5, e6);
  J[7] = -triple3(e1, e5, e6);
  J[8] = -triple3(e4, e2, e6);
  J[9] = -triple3(e4, e5, e3);
  J[10] = triple3(e2, e4, e7);
  J[11] = -triple3(e3, e4, e7);
  J[12] = triple3(e3, e5, e7);
  J[13] = -triple3(e1, e5, e7);
  J[14] = triple3(e1, e6, e7);
  J[15] = -triple3(e2, e6, e7);
  J[16] = 2.0 * triple3(e4, e5, e6);
  J[17] = triple3(e7, e5, e6);
  J[18] = triple3(e4, e7, e6);
  J[19] = triple3(e4, e5, e7);
  g->mVol = 8.0 * (J[0] + (J[7] + J[8] + J[9]) / 3.0);
  /* mean-strain rows: (8 (xi c23 + et c31 + ze c12) + 8/3 (xi c65 + et c46 + ze c54
   * + hst (c51 + c24) + hut (c62 + c35) + hus (c16 + c43))) / mVol, cab = e_a x e_b */
  double c23[3], c31[3], c12[3], c65[3], c46[3], c54[3], c51[3], c24[3], c62[3], c35[3], c16[3], c43[3];
  cross3(e2, e3, c23);
  cross3(e3, e1, c31);
  cross3(e1, e2, c12);
  cross3(e6, e5, c65);
  cross3(e4, e6, c46);
  cross3(e5, e4, c54);
  cross3(e5, e1, c51);
  cross3(e2, e4, c24);
  cross3(e6, e2, c62);
  cross3(e3, e5, c35);
  cross3(e1, e6, c16);
  cross3(e4, e3, c43);
  for (int i = 0; i < 8; ++i) {
    const double xi = shape_vector(V_XI, i), et = shape_vector(V_ET, i), ze = shape_vector(V_ZE, i);
    const double hst = shape_vector(V_HST, i), hut = shape_vector(V_HUT, i), hus = shape_vector(V_HUS, i);
    for (int d = 0; d < 3; ++d) {
      const double lin = c23[d] * xi + c31[d] * et + c12[d] * ze;
      const double quad = c65[d] * xi + c46[d] * et + c54[d] * ze + (c51[d] + c24[d]) * hst + (c62[d] + c35[d]) * hut +
                          (c16[d] + c43[d]) * hus;
      g->dN[i][d] = (8.0 * lin + (8.0 / 3.0) * quad) / g->mVol;
    }
  }
  /* Jmat[d][a] = sum_i x_d(i) dNloc_a(i) = e_{a+1}[d]; Jinv its inverse */
  double Jm[9], Ji[9];
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < 3; ++a) Jm[3 * d + a] = e[a + 1][d];
  if (!invert_n(3, Jm, Ji)) memset(Ji, 0, sizeof Ji);
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) g->Jinv[a][c] = Ji[3 * a + c];
  /* gamma_g = (I - dNmod x^T) h_g */
  static const int hv[4] = {V_HST, V_HUT, V_HUS, V_HSTU};
  for (int m = 0; m < 4; ++m)
    for (int i = 0; i < 8; ++i) {
      double s = 0.0;
      for (int j = 0; j < 8; ++j) {
        double dx = 0.0;
        for (int k = 0; k < 3; ++k) dx += g->dN[i][k] * xl[k][j];
        s += ((i == j ? 1.0 : 0.0) - dx) * shape_vector(hv[m], j);
      }
      g->gam[m][i] = s;
    }
}

/* The integral over the natural cube of xi^a eta^b zeta^c times the Jacobian determinant
 * polynomial, as GetStab's J combinations give it (ssp_brick.hpp:270-289). */
static double cube_integral(const double J[20], int a, int b, int c) {
  const int key = 100 * a + 10 * b + c;
  switch (key) {
    case 200: return 8.0 * (J[0] / 3.0 + J[7] / 5.0 + J[8] / 9.0 + J[9] / 9.0);   /* J0789 */
    case 20: return 8.0 * (J[0] / 3.0 + J[8] / 5.0 + J[7] / 9.0 + J[9] / 9.0);    /* J0879 */
    case 2: return 8.0 * (J[0] / 3.0 + J[9] / 5.0 + J[7] / 9.0 + J[8] / 9.0);     /* J0978 */
    case 11: return 8.0 * (J[4] / 9.0 + J[17] / 27.0);                              /* J417 */
    case 101: return 8.0 * (J[5] / 9.0 + J[18] / 27.0);                             /* J518 */
    case 110: return 8.0 * (J[6] / 9.0 + J[19] / 27.0);                             /* J619 */
    case 102: return 8.0 * (J[1] / 9.0 + J[12] / 15.0 + J[15] / 27.0);              /* J11215 */
    case 120: return 8.0 * (J[1] / 9.0 + J[15] / 15.0 + J[12] / 27.0);              /* J11512 */
    case 12: return 8.0 * (J[2] / 9.0 + J[11] / 15.0 + J[14] / 27.0);               /* J21114 */
    case 210: return 8.0 * (J[2] / 9.0 + J[14] / 15.0 + J[11] / 27.0);              /* J21411 */
    case 21: return 8.0 * (J[3] / 9.0 + J[10] / 15.0 + J[13] / 27.0);               /* J31013 */
    case 201: return 8.0 * (J[3] / 9.0 + J[13] / 15.0 + J[10] / 27.0);              /* J31310 */
    case 220: return 8.0 * (J[0] / 9.0 + J[7] / 15.0 + J[8] / 15.0 + J[9] / 27.0);  /* J789 */
    case 22: return 8.0 * (J[0] / 9.0 + J[8] / 15.0 + J[9] / 15.0 + J[7] / 27.0);   /* J897 */
    case 202: return 8.0 * (J[0] / 9.0 + J[7] / 15.0 + J[9] / 15.0 + J[8] / 27.0);  /* J798 */
    case 211: return 8.0 * (J[4] / 27.0 + 64.0 * J[17] / 45.0);                     /* J174 */
    case 121: return 8.0 * (J[5] / 27.0 + 64.0 * J[18] / 45.0);                     /* J185 */
    case 112: return 8.0 * (J[6] / 27.0 + 64.0 * J[19] / 45.0);                     /* J196 */
    case 111: return 8.0 * J[16] / 27.0;                                            /* J16 */
    default: return 0.0; /* not used by GetStab */
  }
}

/* The natural gradient of hourglass mode m (0 hst, 1 hut, 2 hus, 3 hstu) along direction a as a
 * monomial (exponents of xi, eta, zeta); 0 when that derivative is zero. */
static int mode_gradient(int m, int a, int ex[3]) {
  static const int mono[4][3] = {{1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
  if (!mono[m][a]) return 0;
  for (int k = 0; k < 3; ++k) ex[k] = mono[m][k] - (k == a ? 1 : 0);
  return 1;
}

/* Isotropic block of two gradient directions g, h: C1 g_d h_d + C3 sum_{f != d} g_f h_f on the
 * diagonal, C2 g_d h_e + C3 g_e h_d off it. */
static double iso_entry(const double g[3], const double h[3], double C1, double C2, double C3, int d, int e) {
  if (d == e) {
    double s = 0.0;
    for (int f = 0; f < 3; ++f)
      if (f != d) s += g[f] * h[f];
    return C1 * g[d] * h[d] + C3 * s;
  }
  return C2 * g[d] * h[e] + C3 * g[e] * h[d];
}

/* interior = FCF - KuT FeCFe^-1 FeCFhg (ssp_brick.hpp:265-796), from C1 = C(0,0), C2 = C(0,1),
 * C3 = C(3,3) of the initial tangent. */
static void ssp_interior(const ssp_geom* g, double C1, double C2, double C3, double interior[12][12]) {
  const double* J = g->J;
  /* FCF: P_gh[X][Y] = sum_ab I(d_a h_g d_b h_h) Jinv[a][X] Jinv[b][Y] */
  double FCF[12][12];
  for (int mg = 0; mg < 4; ++mg)
    for (int mh = 0; mh < 4; ++mh) {
      double P[3][3] = {{0}};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          int ea[3], eb[3];
          if (!mode_gradient(mg, a, ea) || !mode_gradient(mh, b, eb)) continue;
          const double I = cube_integral(J, ea[0] + eb[0], ea[1] + eb[1], ea[2] + eb[2]);
          for (int X = 0; X < 3; ++X)
            for (int Y = 0; Y < 3; ++Y) P[X][Y] += I * g->Jinv[a][X] * g->Jinv[b][Y];
        }
      const int shear_only = mg == 3 || mh == 3;
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) {
          double v;
          if (d == e) {
            double s = 0.0;
            for (int f = 0; f < 3; ++f)
              if (f != d) s += P[f][f];
            v = (shear_only ? 0.0 : C1 * P[d][d]) + C3 * s;
          } else {
            v = shear_only ? C3 * P[e][d] : C2 * P[d][e] + C3 * P[e][d];
          }
          FCF[3 * mg + d][3 * mh + e] = v;
        }
    }
  /* enhanced strain: C_ab = iso(Jinv[a], Jinv[b]); FeCFe(a d, b e) = C_ab[d][e] I(xi_a xi_b);
   * FeCFhg(a d, g e) = sum_b C_ab[d][e] I(xi_a d_b h_g), zero for hstu */
  double FeCFe[81], FeCFhg[9][12];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      int ex[3] = {0, 0, 0};
      ex[a] += 1;
      ex[b] += 1;
      const double I = cube_integral(J, ex[0], ex[1], ex[2]);
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) FeCFe[(3 * a + d) * 9 + 3 * b + e] = iso_entry(g->Jinv[a], g->Jinv[b], C1, C2, C3, d, e) * I;
    }
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d)
      for (int m = 0; m < 4; ++m)
        for (int e = 0; e < 3; ++e) {
          double s = 0.0;
          if (m < 3)
            for (int b = 0; b < 3; ++b) {
              int eb[3];
              if (!mode_gradient(m, b, eb)) continue;
              eb[a] += 1;
              s += iso_entry(g->Jinv[a], g->Jinv[b], C1, C2, C3, d, e) * cube_integral(J, eb[0], eb[1], eb[2]);
            }
          FeCFhg[3 * a + d][3 * m + e] = s;
        }
  double inv[81];
  if (!invert_n(9, FeCFe, inv)) memset(inv, 0, sizeof inv);
  double KuTinv[12][9];
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 9; ++j) {
      double s = 0.0;
      for (int k = 0; k < 9; ++k) s += FeCFhg[k][i] * inv[k * 9 + j];
      KuTinv[i][j] = s;
    }
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      double s = 0.0;
      for (int k = 0; k < 9; ++k) s += KuTinv[i][k] * FeCFhg[k][j];
      interior[i][j] = FCF[i][j] - s;
    }
}

/* Node i's volume weight: the integral of N_i (2 x 2 x 2 Gauss, nodal_volumes :800-811) or
 * OpenSees' J[0] * polyJac bracket (poly_bracket :252-259; exact for parallelepipeds only). */
static void ssp_node_volumes(const double xl[3][8], const ssp_geom* g, int opensees_mass, double w[8]) {
  if (!opensees_mass) {
    for (int j = 0; j < 8; ++j) w[j] = 0.0;
    for (int gp = 0; gp < 8; ++gp) {
      double shp[4][8], dvol;
      fxo_brick_shape(xl, gp, shp, &dvol);
      for (int j = 0; j < 8; ++j) w[j] += shp[3][j] * dvol;
    }
    return;
  }
  const double* J = g->J;
  for (int i = 0; i < 8; ++i) {
    const double xi = shape_vector(V_XI, i), et = shape_vector(V_ET, i), ze = shape_vector(V_ZE, i);
    const double hst = shape_vector(V_HST, i), hut = shape_vector(V_HUT, i), hus = shape_vector(V_HUS, i);
    const double hstu = shape_vector(V_HSTU, i);
    const double br = 1.0 + (J[1] * xi + J[2] * et + J[3] * ze + J[7] + J[8] + J[9]) / 3.0 +
                      (J[4] * hut + J[5] * hus + J[6] * hst + J[10] * ze + J[11] * et + J[12] * xi + J[13] * ze +
                       J[14] * et + J[15] * xi) /
                          9.0 +
                      (J[16] * hstu + J[17] * hut + J[18] * hus + J[19] * hst) / 27.0;
    w[i] = J[0] * br;
  }
}

/* geometry and constants with interior from the tangent D (C1 = D00, C2 = D01, C3 = D33) */
static void ssp_element_D(const double xe[8][3], int opensees_mass, const double D[6][6], double* geom,
                          double* consts) {
  double xl[3][8];
  for (int i = 0; i < 8; ++i)
    for (int d = 0; d < 3; ++d) xl[d][i] = xe[i][d];
  ssp_geom g;
  ssp_geometry(xl, &g);
  for (int k = 0; k < 24; ++k) geom[k] = g.dN[k / 3][k % 3];
  geom[24] = g.mVol;
  ssp_node_volumes(xl, &g, opensees_mass, geom + 25);
  double interior[12][12];
  ssp_interior(&g, D[0][0], D[0][1], D[3][3], interior);
  for (int m = 0; m < 4; ++m)
    for (int i = 0; i < 8; ++i) consts[8 * m + i] = g.gam[m][i];
  int k = 32;
  for (int r = 0; r < 12; ++r)
    for (int q = r; q < 12; ++q) consts[k++] = interior[r][q];
}

void fxo_ssp_element(const fxo_elastic* p, const double xe[8][3], int opensees_mass, double* geom, double* consts) {
  double D[6][6];
  fxo_elastic_tangent(p->E, p->nu, D);
  ssp_element_D(xe, opensees_mass, (const double(*)[6])D, geom, consts);
}

/* interior(r, q) from the packed upper triangle (ssp_brick.hpp:917-926) */
static double interior_at(const double* c, int r, int q) {
  if (r > q) {
    const int t = r;
    r = q;
    q = t;
  }
  return c[32 + 12 * r - r * (r - 1) / 2 + (q - r)];
}

/* z = interior (Mben v) (element_term_reduce :964-979) */
static void term_reduce(const double* c, const double* v, double z[12]) {
  double w[12];
  for (int m = 0; m < 4; ++m)
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int i = 0; i < 8; ++i) s += c[8 * m + i] * v[3 * i + d];
      w[3 * m + d] = s;
    }
  for (int r = 0; r < 12; ++r) {
    double s = 0.0;
    for (int q = 0; q < 12; ++q) s += interior_at(c, r, q) * w[q];
    z[r] = s;
  }
}
/* (Kstab v) at node j (element_term_node :981-989) */
static void term_node(const double* c, const double z[12], int j, double out[3]) {
  for (int d = 0; d < 3; ++d) {
    double s = 0.0;
    for (int m = 0; m < 4; ++m) s += c[8 * m + j] * z[3 * m + d];
    out[d] = s;
  }
}
/* Kstab(a, b), in the order of (min, max) (element_term_entry :991-1005) */
static double term_entry(const double* c, int a, int b) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  const int i = a / 3, d = a % 3, k = b / 3, e = b % 3;
  double s = 0.0;
  for (int g = 0; g < 4; ++g)
    for (int h = 0; h < 4; ++h) s += c[8 * g + i] * interior_at(c, 3 * g + d, 3 * h + e) * c[8 * h + k];
  return s;
}

static void mean_strain(const double* geom, const double* u, double eps[6]) {
  for (int c = 0; c < 6; ++c) eps[c] = 0.0;
  for (int i = 0; i < 8; ++i) {
    const double bx = geom[3 * i], by = geom[3 * i + 1], bz = geom[3 * i + 2];
    const double u0 = u[3 * i], u1 = u[3 * i + 1], u2 = u[3 * i + 2];
    eps[0] += bx * u0;
    eps[3] += by * u0;
    eps[5] += bz * u0;
    eps[1] += by * u1;
    eps[3] += bx * u1;
    eps[4] += bz * u1;
    eps[2] += bz * u2;
    eps[4] += by * u2;
    eps[5] += bx * u2;
  }
}
/* mVol Bnot_j^T sigma (point_node_force :1035-1041) */
static void node_force(const double* geom, const double sig[6], int j, double f[3]) {
  const double bx = geom[3 * j], by = geom[3 * j + 1], bz = geom[3 * j + 2], V = geom[24];
  f[0] = (bx * sig[0] + by * sig[3] + bz * sig[5]) * V;
  f[1] = (by * sig[1] + bx * sig[3] + bz * sig[4]) * V;
  f[2] = (bz * sig[2] + by * sig[4] + bx * sig[5]) * V;
}
/* mVol Bnot_J^T D Bnot_K (point_block_tangent :1049-1067) */
static void block_tangent(const double* geom, const double* D, int J, int K, double out[3][3]) {
  double BJ[6][3] = {{0}}, BK[6][3] = {{0}};
  const int nodes[2] = {J, K};
  for (int t = 0; t < 2; ++t) {
    double(*B)[3] = t == 0 ? BJ : BK;
    const double* b = geom + 3 * nodes[t];
    B[0][0] = b[0];
    B[1][1] = b[1];
    B[2][2] = b[2];
    B[3][0] = b[1];
    B[3][1] = b[0];
    B[4][1] = b[2];
    B[4][2] = b[1];
    B[5][0] = b[2];
    B[5][2] = b[0];
  }
  double BD[3][6];
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 6; ++q) {
      double s = 0.0;
      for (int r = 0; r < 6; ++r) s += BJ[r][p] * (D[6 * r + q] * geom[24]);
      BD[p][q] = s;
    }
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      double s = 0.0;
      for (int r = 0; r < 6; ++r) s += BD[p][r] * BK[r][q];
      out[p][q] = s;
    }
}

static void element_data(const fxo_elastic* p, const double xe[8][3], int opensees_mass, double geom[33],
                         double consts[110], double D[36]) {
  fxo_ssp_element(p, xe, opensees_mass, geom, consts);
  fxo_elastic_tangent(p->E, p->nu, (double(*)[6])D);
}

void fxo_ssp_force(const fxo_elastic* p, const double xe[8][3], int opensees_mass, const double* ue, const double* ve,
                   double* fe, double* Ke) {
  double geom[33], c[110], D[36];
  element_data(p, xe, opensees_mass, geom, c, D);
  const fxo_rayleigh* r = &p->damp;
  double zu[12], zv[12];
  term_reduce(c, ue, zu);
  term_reduce(c, ve, zv);
  const double beta = r->betaK + r->betaKinit + r->betaKcomm;
  double resid[24], fdamp[24];
  for (int j = 0; j < 8; ++j) {
    term_node(c, zu, j, resid + 3 * j);
    if (beta != 0.0) {
      double kv[3];
      term_node(c, zv, j, kv);
      for (int q = 0; q < 3; ++q) fdamp[3 * j + q] = beta * kv[q];
    } else {
      for (int q = 0; q < 3; ++q) fdamp[3 * j + q] = 0.0;
    }
  }
  if (Ke)
    for (int a = 0; a < 24; ++a)
      for (int b = 0; b < 24; ++b) Ke[24 * a + b] = term_entry(c, a, b);
  double eps[6], sig[6], dsig[6] = {0, 0, 0, 0, 0, 0};
  mean_strain(geom, ue, eps);
  for (int a = 0; a < 6; ++a) {
    double s = 0.0;
    for (int q = 0; q < 6; ++q) s += D[6 * a + q] * eps[q];
    sig[a] = s;
  }
  const int stiff_damp = r->betaK != 0.0 || r->betaKinit != 0.0 || r->betaKcomm != 0.0;
  if (stiff_damp) {
    double rate[6];
    mean_strain(geom, ve, rate);
    for (int a = 0; a < 6; ++a) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) {
        double Daq = r->betaK * D[6 * a + q];
        if (r->betaKinit != 0.0) Daq += r->betaKinit * D[6 * a + q];
        if (r->betaKcomm != 0.0) Daq += r->betaKcomm * D[6 * a + q];
        s += Daq * rate[q];
      }
      dsig[a] = s;
    }
  }
  for (int j = 0; j < 8; ++j) {
    double fj[3], dj[3] = {0, 0, 0};
    node_force(geom, sig, j, fj);
    if (stiff_damp) node_force(geom, dsig, j, dj);
    const double m = p->rho * geom[25 + j];
    for (int q = 0; q < 3; ++q) {
      if (r->alphaM != 0.0) dj[q] += r->alphaM * m * ve[3 * j + q];
      resid[3 * j + q] += fj[q];
      resid[3 * j + q] -= p->b[q] * geom[25 + j];
      fdamp[3 * j + q] += dj[q];
    }
  }
  if (Ke)
    for (int J = 0; J < 8; ++J)
      for (int K = 0; K < 8; ++K) {
        double blk[3][3];
        block_tangent(geom, D, J, K, blk);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) Ke[24 * (3 * J + a) + 3 * K + b] += blk[a][b];
      }
  for (int a = 0; a < 24; ++a) fe[a] = resid[a] + fdamp[a];
}

void fxo_ssp_mass(const fxo_elastic* p, const double xe[8][3], int opensees_mass, double* Me) {
  double geom[33], c[110];
  fxo_ssp_element(p, xe, opensees_mass, geom, c);
  memset(Me, 0, sizeof(double) * 576);
  for (int i = 0; i < 8; ++i)
    for (int q = 0; q < 3; ++q) Me[(3 * i + q) * 24 + 3 * i + q] = p->rho * geom[25 + i];
}

/* element_operator.hpp Factors (:45-56) with initial_stiffness_is_tangent: cK folds betaKinit
 * in, c0 = 0; the element term's factor cE (element_term.hpp:34-37). */
static void ssp_factors(const fxo_rayleigh* d, const fxo_coeffs* k, double* cK, double* cc, double* cM, double* cE) {
  *cK = k->stiffness + k->damping * d->betaK + k->damping * d->betaKinit;
  *cc = k->damping * d->betaKcomm;
  *cM = k->mass + k->damping * d->alphaM;
  *cE = k->stiffness + k->damping * (d->betaK + d->betaKinit + d->betaKcomm);
}

void fxo_ssp_operator_apply(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3], const double* ve,
                            double* ye) {
  double geom[33], c[110], D[36];
  element_data(p, xe, 0, geom, c, D);
  double cK, cc, cM, cE;
  ssp_factors(&p->damp, k, &cK, &cc, &cM, &cE);
  double Deff[36];
  for (int a = 0; a < 36; ++a) {
    double v = cK * D[a];
    if (cc != 0.0) v += cc * D[a];
    Deff[a] = v;
  }
  double eps[6], sig[6];
  mean_strain(geom, ve, eps);
  for (int a = 0; a < 6; ++a) {
    double s = 0.0;
    for (int q = 0; q < 6; ++q) s += Deff[6 * a + q] * eps[q];
    sig[a] = s;
  }
  for (int j = 0; j < 8; ++j) {
    double fj[3];
    node_force(geom, sig, j, fj);
    for (int q = 0; q < 3; ++q) {
      ye[3 * j + q] = fj[q];
      if (cM != 0.0) ye[3 * j + q] += cM * (p->rho * geom[25 + j] * ve[3 * j + q]);
    }
  }
  if (cE != 0.0) {
    double z[12];
    term_reduce(c, ve, z);
    for (int j = 0; j < 8; ++j) {
      double kj[3];
      term_node(c, z, j, kj);
      for (int q = 0; q < 3; ++q) ye[3 * j + q] += cE * kj[q];
    }
  }
}

void fxo_ssp_operator_diagonal(const fxo_elastic* p, const fxo_coeffs* k, const double xe[8][3], const int32_t* eq,
                               double* de) {
  double geom[33], c[110], D[36];
  element_data(p, xe, 0, geom, c, D);
  double cK, cc, cM, cE;
  ssp_factors(&p->damp, k, &cK, &cc, &cM, &cE);
  double Deff[36];
  for (int a = 0; a < 36; ++a) {
    double v = cK * D[a];
    if (cc != 0.0) v += cc * D[a];
    Deff[a] = v;
  }
  int first[24], merged = 0;
  for (int a = 0; a < 24; ++a) {
    de[a] = 0.0;
    first[a] = a;
    if (eq[a] < 0) continue;
    for (int b = 0; b < a; ++b)
      if (eq[b] == eq[a]) {
        first[a] = first[b];
        merged = 1;
        break;
      }
  }
  for (int j = 0; j < 8; ++j) {
    double blk[3][3];
    block_tangent(geom, Deff, j, j, blk);
    const double m = cM != 0.0 ? cM * (p->rho * geom[25 + j]) : 0.0;
    for (int q = 0; q < 3; ++q) de[3 * j + q] += blk[q][q] + m;
  }
  if (merged)
    for (int J = 0; J < 8; ++J)
      for (int K = 0; K < 8; ++K) {
        double blk[3][3];
        block_tangent(geom, Deff, J, K, blk);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const int ia = 3 * J + a, ib = 3 * K + b;
            if (ia == ib || eq[ia] < 0 || eq[ia] != eq[ib]) continue;
            de[first[ia]] += blk[a][b]; /* the lumped mass has no off-diagonal term */
          }
      }
  if (cE != 0.0) {
    for (int a = 0; a < 24; ++a) de[a] += cE * term_entry(c, a, a);
    if (merged)
      for (int a = 0; a < 24; ++a)
        for (int b = 0; b < 24; ++b) {
          if (a == b || eq[a] < 0 || eq[a] != eq[b]) continue;
          de[first[a]] += cE * term_entry(c, a, b);
        }
  }
}

double fxo_ssp_max_omega_sq(const fxo_elastic* p, const double xe[8][3]) {
  double geom[33], c[110], D[36];
  element_data(p, xe, 0, geom, c, D);
  double A[24][24];
  for (int a = 0; a < 24; ++a)
    for (int b = 0; b < 24; ++b) A[a][b] = term_entry(c, a, b);
  for (int J = 0; J < 8; ++J)
    for (int K = 0; K < 8; ++K) {
      double blk[3][3];
      block_tangent(geom, D, J, K, blk);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) A[3 * J + a][3 * K + b] += blk[a][b];
    }
  double s[8];
  for (int j = 0; j < 8; ++j) {
    const double m = p->rho * geom[25 + j];
    if (!(m > 0.0)) return 0.0;
    s[j] = 1.0 / sqrt(m);
  }
  for (int a = 0; a < 24; ++a)
    for (int b = 0; b < 24; ++b) A[a][b] *= s[a / 3] * s[b / 3];
  return fxo_max_eigenvalue_symmetric24(A);
}

/* ---- SSPbrick<PressureIndependMultiYield> ---------------------------------------------------- */
static void pimy_element_data(const fxo_pimy_elem* p, const double xe[8][3], double geom[33], double c[110]) {
  double D0[36];
  fxo_pimy_initial_tangent(&p->mat, D0);
  ssp_element_D(xe, 0, (const double(*)[6])D0, geom, c);
}

void fxo_ssp_force_pimy(const fxo_pimy_elem* p, const double xe[8][3], const double* ue, const double* ve,
                        const double* C0, double* T0, double* fe, double* Dout) {
  double geom[33], c[110];
  pimy_element_data(p, xe, geom, c);
  const fxo_rayleigh* r = &p->damp;
  const int stiff = r->betaK != 0.0 || r->betaKinit != 0.0 || r->betaKcomm != 0.0;
  const int damped = ve && (stiff || r->alphaM != 0.0);
  double zu[12], zv[12];
  term_reduce(c, ue, zu);
  if (damped) term_reduce(c, ve, zv);
  const double beta = r->betaK + r->betaKinit + r->betaKcomm;
  double resid[24], fdamp[24];
  for (int j = 0; j < 8; ++j) {
    term_node(c, zu, j, resid + 3 * j);
    if (damped && beta != 0.0) {
      double kv[3];
      term_node(c, zv, j, kv);
      for (int q = 0; q < 3; ++q) fdamp[3 * j + q] = beta * kv[q];
    } else {
      for (int q = 0; q < 3; ++q) fdamp[3 * j + q] = 0.0;
    }
  }
  double eps[6], sig[6], D[36], dsig[6] = {0, 0, 0, 0, 0, 0};
  mean_strain(geom, ue, eps);
  fxo_pimy_update(&p->mat, C0, T0, eps, sig, D);
  if (Dout) memcpy(Dout, D, sizeof D);
  if (damped && stiff) {
    double Dc[36], rate[6];
    if (r->betaKcomm != 0.0) fxo_pimy_committed_tangent(&p->mat, C0, Dc);
    mean_strain(geom, ve, rate);
    for (int a = 0; a < 6; ++a) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) {
        double Daq = r->betaK * D[6 * a + q];
        if (r->betaKinit != 0.0) Daq += r->betaKinit * D[6 * a + q]; /* the initial stiffness is the tangent */
        if (r->betaKcomm != 0.0) Daq += r->betaKcomm * Dc[6 * a + q];
        s += Daq * rate[q];
      }
      dsig[a] = s;
    }
  }
  for (int j = 0; j < 8; ++j) {
    double fj[3], dj[3] = {0, 0, 0};
    node_force(geom, sig, j, fj);
    if (damped) {
      if (stiff) node_force(geom, dsig, j, dj);
      if (r->alphaM != 0.0) {
        const double m = p->mat.rho * geom[25 + j];
        for (int q = 0; q < 3; ++q) dj[q] += r->alphaM * m * ve[3 * j + q];
      }
    }
    for (int q = 0; q < 3; ++q) {
      resid[3 * j + q] += fj[q];
      resid[3 * j + q] -= p->b[q] * geom[25 + j];
      fdamp[3 * j + q] += dj[q];
    }
  }
  for (int a = 0; a < 24; ++a) fe[a] = damped ? resid[a] + fdamp[a] : resid[a];
}

/* D_eff = cK Dt + cc Dc (c0 = 0: the initial stiffness is the tangent) */
static void pimy_ssp_effective(const fxo_pimy_elem* p, double cK, double cc, const double* Dt, const double* C0,
                               double* Deff) {
  double Dc[36];
  if (cc != 0.0) fxo_pimy_committed_tangent(&p->mat, C0, Dc);
  for (int a = 0; a < 36; ++a) {
    double v = cK * Dt[a];
    if (cc != 0.0) v += cc * Dc[a];
    Deff[a] = v;
  }
}

void fxo_ssp_operator_apply_pimy(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3], const double* Dt,
                                 const double* C0, const double* ve, double* ye) {
  double geom[33], c[110];
  pimy_element_data(p, xe, geom, c);
  double cK, cc, cM, cE, Deff[36];
  ssp_factors(&p->damp, k, &cK, &cc, &cM, &cE);
  pimy_ssp_effective(p, cK, cc, Dt, C0, Deff);
  double eps[6], sig[6];
  mean_strain(geom, ve, eps);
  for (int a = 0; a < 6; ++a) {
    double s = 0.0;
    for (int q = 0; q < 6; ++q) s += Deff[6 * a + q] * eps[q];
    sig[a] = s;
  }
  for (int j = 0; j < 8; ++j) {
    double fj[3];
    node_force(geom, sig, j, fj);
    for (int q = 0; q < 3; ++q) {
      ye[3 * j + q] = fj[q];
      if (cM != 0.0) ye[3 * j + q] += cM * (p->mat.rho * geom[25 + j] * ve[3 * j + q]);
    }
  }
  if (cE != 0.0) {
    double z[12];
    term_reduce(c, ve, z);
    for (int j = 0; j < 8; ++j) {
      double kj[3];
      term_node(c, z, j, kj);
      for (int q = 0; q < 3; ++q) ye[3 * j + q] += cE * kj[q];
    }
  }
}

void fxo_ssp_operator_diagonal_pimy(const fxo_pimy_elem* p, const fxo_coeffs* k, const double xe[8][3],
                                    const double* Dt, const double* C0, const int32_t* eq, double* de) {
  double geom[33], c[110];
  pimy_element_data(p, xe, geom, c);
  double cK, cc, cM, cE, Deff[36];
  ssp_factors(&p->damp, k, &cK, &cc, &cM, &cE);
  pimy_ssp_effective(p, cK, cc, Dt, C0, Deff);
  int first[24], merged = 0;
  for (int a = 0; a < 24; ++a) {
    de[a] = 0.0;
    first[a] = a;
    if (eq[a] < 0) continue;
    for (int b = 0; b < a; ++b)
      if (eq[b] == eq[a]) {
        first[a] = first[b];
        merged = 1;
        break;
      }
  }
  for (int j = 0; j < 8; ++j) {
    double blk[3][3];
    block_tangent(geom, Deff, j, j, blk);
    const double m = cM != 0.0 ? cM * (p->mat.rho * geom[25 + j]) : 0.0;
    for (int q = 0; q < 3; ++q) de[3 * j + q] += blk[q][q] + m;
  }
  if (merged)
    for (int J = 0; J < 8; ++J)
      for (int K = 0; K < 8; ++K) {
        double blk[3][3];
        block_tangent(geom, Deff, J, K, blk);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const int ia = 3 * J + a, ib = 3 * K + b;
            if (ia == ib || eq[ia] < 0 || eq[ia] != eq[ib]) continue;
            de[first[ia]] += blk[a][b];
          }
      }
  if (cE != 0.0) {
    for (int a = 0; a < 24; ++a) de[a] += cE * term_entry(c, a, a);
    if (merged)
      for (int a = 0; a < 24; ++a)
        for (int b = 0; b < 24; ++b) {
          if (a == b || eq[a] < 0 || eq[a] != eq[b]) continue;
          de[first[a]] += cE * term_entry(c, a, b);
        }
  }
}
