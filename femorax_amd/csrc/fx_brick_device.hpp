// stdBrick shape functions at the 2 x 2 x 2 Gauss points, for the device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fx {

// Natural signs of the 8 nodes (OpenSees Brick order).
static __device__ __constant__ int8_t kSx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
static __device__ __constant__ int8_t kSy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static __device__ __constant__ int8_t kSz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

template <typename T>
struct Shape {
  T dx[8], dy[8], dz[8], N[8];
  T dvol;
};

// shp3d at Gauss point gp = 4i + 2j + k (std_brick.hpp:45-152, 203-211).
template <typename T>
__device__ __forceinline__ void gauss_shape(const T (&x)[8], const T (&y)[8], const T (&z)[8], int gp, Shape<T>& s) {
  const T r3 = T(0.57735026918962576451);  // 1/sqrt(3)
  const T xi = (gp & 4) ? r3 : -r3, eta = (gp & 2) ? r3 : -r3, zeta = (gp & 1) ? r3 : -r3;
  T nx[8], ny[8], nz[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const T fx = kSx[k] > 0 ? T(1) + xi : T(1) - xi;
    const T fy = kSy[k] > 0 ? T(1) + eta : T(1) - eta;
    const T fz = kSz[k] > 0 ? T(1) + zeta : T(1) - zeta;
    const T cyz = T(0.125) * fy * fz, cxz = T(0.125) * fx * fz, cxy = T(0.125) * fx * fy;
    nx[k] = kSx[k] > 0 ? cyz : -cyz;
    ny[k] = kSy[k] > 0 ? cxz : -cxz;
    nz[k] = kSz[k] > 0 ? cxy : -cxy;
    s.N[k] = cxy * fz;
  }
  T J[3][3];
  const T* c[3] = {x, y, z};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const T* p = c[j];
    J[j][0] = (p[1] - p[0]) * nx[1] + (p[2] - p[3]) * nx[2] + (p[5] - p[4]) * nx[5] + (p[6] - p[7]) * nx[6];
    J[j][1] = (p[2] - p[1]) * ny[2] + (p[3] - p[0]) * ny[3] + (p[6] - p[5]) * ny[6] + (p[7] - p[4]) * ny[7];
    J[j][2] = (p[4] - p[0]) * nz[4] + (p[5] - p[1]) * nz[5] + (p[6] - p[2]) * nz[6] + (p[7] - p[3]) * nz[7];
  }
  T ad[3][3];
  ad[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  ad[0][1] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
  ad[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  ad[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  ad[1][1] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
  ad[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  ad[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  ad[2][1] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
  ad[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const T det = J[0][0] * ad[0][0] + J[0][1] * ad[1][0] + J[0][2] * ad[2][0];
  const T rdet = T(1) / det;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const T a = nx[k], b = ny[k], cc = nz[k];
    s.dx[k] = (a * ad[0][0] + b * ad[1][0] + cc * ad[2][0]) * rdet;
    s.dy[k] = (a * ad[0][1] + b * ad[1][1] + cc * ad[2][1]) * rdet;
    s.dz[k] = (a * ad[0][2] + b * ad[1][2] + cc * ad[2][2]) * rdet;
  }
  s.dvol = det;
}

}  // namespace fx

namespace fx {

// The stdBrick operator y_e = sum_gp B^T (D_eff B v) dvol + cM' (M v) by sum factorisation on the
// trilinear modes (the element is the same, the arithmetic is rearranged; element_operator_apply,
// element_operator.hpp:82-130; shp3d at the 2 x 2 x 2 Gauss points, std_brick.hpp:45-152).
// With corner values in lattice order q = i + 2j + 4k, the modal coefficients a = H v / 8 give the
// field as sum_m a_m xi^(m&1) eta^(m>>1&1) zeta^(m>>2&1), so the natural gradients of the
// displacement and of the coordinates at a Gauss point take four terms each; the natural flux
// P = dvol sigma J^-T goes back through the modes' gradients into the modal force G, and the node
// forces are H^T G / 8. About 1,900 FP64 operations per element where the shape-function form
// takes about 3,500.
// Corner q of the lattice order holds the reference's node kNodeOfCorner[q].
__host__ __device__ constexpr int node_of_corner(int q) { return (q & 2) ? ((q & 1) ? q - 1 : q + 1) : q; }

template <typename T>
__host__ __device__ __forceinline__ void hadamard8_hd(T (&v)[8]) {
#pragma unroll
  for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (!(q & bit)) {
        const T a = v[q], b = v[q | bit];
        v[q] = a + b;
        v[q | bit] = b - a;
      }
}

template <typename T>
__host__ __device__ __forceinline__ void hadamard8t_hd(T (&g)[8]) {
#pragma unroll
  for (int bit = 4; bit >= 1; bit >>= 1)
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (!(q & bit)) {
        const T a = g[q], b = g[q | bit];
        g[q] = a - b;
        g[q | bit] = a + b;
      }
}

// x, y, z: node coordinates, v: node values (node-major, 3 per node), reference node order.
// lam, mu: the effective D's Lame factors; rhoM: cM' rho (0: no mass term); consistent: the
// consistent mass (else lumped, sum_gp N_j dvol). vl(i) returns v[i] again for the lumped mass
// term at the end (a caller whose values sit in LDS reloads them there instead of keeping 24
// registers live through the Gauss loop).
template <typename T, typename VL>
__host__ __device__ __forceinline__ void element_apply_modal(const T (&x)[8], const T (&y)[8], const T (&z)[8],
                                                             const T (&v)[24], T lam, T mu, T rhoM, bool consistent,
                                                             T (&out)[24], VL vl) {
  T c[3][8], a[3][8], G[3][8], W[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int n = node_of_corner(q);
    c[0][q] = x[n];
    c[1][q] = y[n];
    c[2][q] = z[n];
#pragma unroll
    for (int d = 0; d < 3; ++d) a[d][q] = v[3 * n + d];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    hadamard8_hd(c[d]);
    hadamard8_hd(a[d]);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      c[d][m] *= T(0.125);
      a[d][m] *= T(0.125);
      G[d][m] = T(0);
    }
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) W[m] = T(0);
  const T g = T(0.57735026918962576451);  // 1/sqrt(3)
  const T m2 = mu + mu;
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    const T xi = (gp & 4) ? g : -g, et = (gp & 2) ? g : -g, ze = (gp & 1) ? g : -g;
    const T xe = xi * et, xz = xi * ze, ez = et * ze;
    // natural gradients: d/dxi, d/deta, d/dzeta of a modal field f
    auto grad = [&](const T (&f)[8], T (&o)[3]) {
      o[0] = f[1] + f[3] * et + f[5] * ze + f[7] * ez;
      o[1] = f[2] + f[3] * xi + f[6] * ze + f[7] * xz;
      o[2] = f[4] + f[5] * xi + f[6] * et + f[7] * xe;
    };
    T J[3][3], Gn[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      grad(c[d], J[d]);
      grad(a[d], Gn[d]);
    }
    T ad[3][3];
    ad[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    ad[0][1] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    ad[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    ad[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    ad[1][1] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    ad[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    ad[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    ad[2][1] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    ad[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const T det = J[0][0] * ad[0][0] + J[0][1] * ad[1][0] + J[0][2] * ad[2][0];
    const T rdet = T(1) / det;
    // physical gradient H[c][e] = sum_a Gn[c][a] Jinv[a][e], Jinv = ad / det
    T H[3][3];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int e = 0; e < 3; ++e)
        H[cc][e] = (Gn[cc][0] * ad[0][e] + Gn[cc][1] * ad[1][e] + Gn[cc][2] * ad[2][e]) * rdet;
    const T lt = lam * (H[0][0] + H[1][1] + H[2][2]);
    T S[3][3];
    S[0][0] = lt + m2 * H[0][0];
    S[1][1] = lt + m2 * H[1][1];
    S[2][2] = lt + m2 * H[2][2];
    S[0][1] = S[1][0] = mu * (H[0][1] + H[1][0]);
    S[1][2] = S[2][1] = mu * (H[1][2] + H[2][1]);
    S[0][2] = S[2][0] = mu * (H[0][2] + H[2][0]);
    // natural flux P[c][a] = dvol sum_e S[c][e] Jinv[a][e] = sum_e S[c][e] ad[a][e] (dvol = det)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      T P[3];
#pragma unroll
      for (int aa = 0; aa < 3; ++aa) P[aa] = S[cc][0] * ad[aa][0] + S[cc][1] * ad[aa][1] + S[cc][2] * ad[aa][2];
      G[cc][1] += P[0];
      G[cc][2] += P[1];
      G[cc][4] += P[2];
      G[cc][3] += et * P[0] + xi * P[1];
      G[cc][5] += ze * P[0] + xi * P[2];
      G[cc][6] += ze * P[1] + et * P[2];
      G[cc][7] += ez * P[0] + xz * P[1] + xe * P[2];
    }
    if (rhoM != T(0)) {
      const T phi[8] = {T(1), xi, et, xe, ze, xz, ez, xe * ze};
      if (consistent) {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          T u = T(0);
#pragma unroll
          for (int m = 0; m < 8; ++m) u += a[cc][m] * phi[m];
          const T t = rhoM * det * u;
#pragma unroll
          for (int m = 0; m < 8; ++m) G[cc][m] += phi[m] * t;
        }
      } else {
#pragma unroll
        for (int m = 0; m < 8; ++m) W[m] += phi[m] * det;
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) hadamard8t_hd(G[d]);
  if (rhoM != T(0) && !consistent) hadamard8t_hd(W);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int n = node_of_corner(q);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T o = T(0.125) * G[d][q];
      if (rhoM != T(0) && !consistent) o += rhoM * (T(0.125) * W[q]) * vl(3 * n + d);
      out[3 * n + d] = o;
    }
  }
}

template <typename T>
__host__ __device__ __forceinline__ void element_apply_modal(const T (&x)[8], const T (&y)[8], const T (&z)[8],
                                                             const T (&v)[24], T lam, T mu, T rhoM, bool consistent,
                                                             T (&out)[24]) {
  element_apply_modal<T>(x, y, z, v, lam, mu, rhoM, consistent, out, [&](int i) { return v[i]; });
}

}  // namespace fx
