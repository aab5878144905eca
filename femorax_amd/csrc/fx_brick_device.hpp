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
