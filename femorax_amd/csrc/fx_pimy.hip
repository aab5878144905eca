// PressureIndependMultiYield device return map (the algorithm and its sources are listed in
// fx_pimy.hpp), the material-point driver, and the stdBrick<PressureIndependMultiYield> batch
// kernels: internal force with the return map at every Gauss point, the operator and its
// diagonal on the stored tangents.
//
// The batch kernels run one thread per Gauss point (8 consecutive lanes per element), so the
// return maps of an element run side by side and the SoA history is read coalesced. Each point
// leaves its node contributions in LDS; the element's first lane adds them in Gauss-point
// order and writes the element's 24 values to its slots; a gather adds the slots of each
// equation in (element, DOF) order (assembly.cpp:64-95). Deterministic.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fx_brick_device.hpp"
#include "fx_internal.hpp"
#include "fx_kernels.hpp"
#include "fx_pimy.hpp"
#include "fx_pimy_point.hpp"

namespace fx {

// Material-point driver: point i uses parameter set param[i], committed history C, trial
// history T (SoA, stride npts), strain eps[k * npts + i]; writes sig[k * npts + i] and the
// row-major tangent D[k * npts + i] (k < 36).
__global__ __launch_bounds__(128) void pimy_point_kernel(int64_t npts, const PimyParams* __restrict__ params,
                                                         const int32_t* __restrict__ param, const double* C, double* T,
                                                         const double* __restrict__ eps, double* __restrict__ sig,
                                                         double* __restrict__ D) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  double e[6], s[6], d[36];
  for (int k = 0; k < 6; ++k) e[k] = eps[k * npts + i];
  pimy_update(params[param[i]], C + i, T + i, npts, e, s, d);
  for (int k = 0; k < 6; ++k) sig[k * npts + i] = s[k];
  if (D)
    for (int k = 0; k < 36; ++k) D[k * npts + i] = d[k];
}

// State at rest: zeros, the elastic moduli and the backbone surfaces (initial_state :142-150).
__global__ void pimy_init_kernel(int64_t npts, const PimyParams* __restrict__ params, const int32_t* __restrict__ param,
                                 double* S) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const PimyParams& p = params[param[i]];
  for (int k = 0; k < kPimyNState; ++k) S[k * npts + i] = 0.0;
  S[kPG * npts + i] = p.G;
  S[kPK * npts + i] = p.K;
  for (int j = 1; j <= kPimyNS; ++j) {
    S[(kPSizes + j - 1) * npts + i] = p.size[j];
    S[(kPModuli + j - 1) * npts + i] = p.modulus[j];
  }
}

// ---- stdBrick<PressureIndependMultiYield> batches ------------------------------------------------

namespace {

constexpr int kPointsPerBlock = 128;  // 16 elements
// batches below this many elements run the operator one thread per point (enough threads to
// fill the device), larger ones one thread per element (no LDS, sums in registers)
constexpr int kPointOperatorBelow = 32768;
constexpr int kElemsPerBlock = kPointsPerBlock / 8;

struct PimyBatchArgs {
  int32_t count, padded, nnode;
  int64_t npts;
  const int32_t* node;
  const int32_t* param;
  const int32_t* node_eq;
  const double* coords;
  const PimyParams* params;
  const double* C;
  double* T;
  double* tangent;  // [21 * npts] upper triangle row-major, or null
  const double* x;  // u (force) or v (operator)
  const double* v;  // force: the velocities of the Rayleigh damping force, or null
  int32_t* touched; // per point: the highest surface whose centre an update has written
  double* slots;    // [24 * padded]
  fx_coeffs k;
  double* cres;     // compacted return map: [33 * npts] stress, tangent (upper), strain
  int32_t* clist;
  int32_t* ccount;
  int32_t unit_dir;  // operator: >= 0 replaces x by the rigid unit translation in that direction
  // SSPbrick batches: the elements' constants c[k * ssp_stride + ssp_pos[e]] (dNmod, mVol, node
  // volume weights, hourglass vectors, interior's upper triangle), Kstab from the material's initial
  // tangent
  const double* ssp;
  int64_t ssp_stride;
  const int32_t* ssp_pos;
};

// The operator's input at local DOF k: x at the DOF's equation (0 at a fixed DOF), or with unit_dir
// >= 0 the rigid unit translation at every DOF, fixed ones included (operator_kernels.hpp:96-98).
__device__ inline double operator_input(const PimyBatchArgs& a, const int (&eq)[24], int k) {
  if (a.unit_dir >= 0) return k % 3 == a.unit_dir ? 1.0 : 0.0;
  return eq[k] >= 0 ? a.x[eq[k]] : 0.0;
}

__device__ inline void element_geometry(const PimyBatchArgs& a, int e, double (&X)[8], double (&Y)[8], double (&Z)[8],
                                        int (&eq)[24]) {
  const size_t P = static_cast<size_t>(a.padded);
  for (int k = 0; k < 8; ++k) {
    const int n = a.node[k * P + e];
    X[k] = a.coords[n];
    Y[k] = a.coords[static_cast<size_t>(a.nnode) + n];
    Z[k] = a.coords[2 * static_cast<size_t>(a.nnode) + n];
    for (int d = 0; d < 3; ++d) eq[3 * k + d] = a.node_eq[3 * n + d];
  }
}

// strain 11 22 33 12 23 31 with engineering shear (std_brick.hpp:276-301)
__device__ inline void point_strain(const Shape<double>& s, const double (&v)[24], double (&eps)[6]) {
  for (int c = 0; c < 6; ++c) eps[c] = 0.0;
  for (int j = 0; j < 8; ++j) {
    const double dx = s.dx[j], dy = s.dy[j], dz = s.dz[j];
    const double u0 = v[3 * j], u1 = v[3 * j + 1], u2 = v[3 * j + 2];
    eps[0] += dx * u0;
    eps[1] += dy * u1;
    eps[2] += dz * u2;
    eps[3] += dy * u0 + dx * u1;
    eps[4] += dz * u1 + dy * u2;
    eps[5] += dz * u0 + dx * u2;
  }
}

// B_j^T (sigma dvol) (std_brick.hpp:305-327)
__device__ inline void node_force(const Shape<double>& s, const double (&sig)[6], int j, double (&out)[3]) {
  double t[6];
  for (int c = 0; c < 6; ++c) t[c] = sig[c] * s.dvol;
  const double dx = s.dx[j], dy = s.dy[j], dz = s.dz[j];
  out[0] = dx * t[0] + dy * t[3] + dz * t[5];
  out[1] = dy * t[1] + dx * t[3] + dz * t[4];
  out[2] = dz * t[2] + dy * t[4] + dx * t[5];
}

// entry (i, j) of the elastic initial tangent of the reference moduli (initial_D)
__device__ inline double initial_entry(const PimyParams& p, int i, int j) {
  double v = 0.;
  if (i == j) v += p.G;
  if (i < 3 && j < 3 && i == j) v += p.G;
  if (i < 3 && j < 3) v += (p.K - 2. * p.G / 3.);
  return v;
}

__device__ inline void initial_D(const PimyParams& p, double (&D)[36]) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double v = 0.;
      if (i == j) v += p.G;
      if (i < 3 && j < 3 && i == j) v += p.G;
      if (i < 3 && j < 3) v += (p.K - 2. * p.G / 3.);
      D[6 * i + j] = v;
    }
}

// The operator factors (element_operator.hpp:45-56; the PIMY initial stiffness is not its
// tangent): cK on the stored tangent, c0 on the initial, cc on the committed tangent, cM on the
// mass.
__device__ inline void operator_factors(const PimyParams& p, const fx_coeffs& k, double& cK, double& c0, double& cc,
                                        double& cM) {
  cK = k.initial ? k.damping * p.betaK : k.stiffness + k.damping * p.betaK;
  c0 = k.initial ? k.stiffness + k.damping * p.betaKinit : k.damping * p.betaKinit;
  cc = k.damping * p.betaKcomm;
  cM = k.mass + k.damping * p.alphaM;
}

// The stored tangents: the 21 upper-triangle entries of each point, entry w of point gp of
// element e at [(w * 8 + gp) * count + e], so that threads on consecutive elements read
// consecutive addresses.
__device__ inline int64_t tangent_at(const PimyBatchArgs& a, int64_t e, int gp, int w) {
  return (static_cast<int64_t>(w) * 8 + gp) * a.count + e;
}

// D_eff = cK D_t + c0 D_0 + cc D_c (effective_tangent, element_operator.hpp:58-75); D_t the stored
// (symmetric) tangent, D_c the committed tangent of the point's committed history
__device__ inline void effective_D(const PimyBatchArgs& a, const PimyParams& p, int64_t e, int gp, double (&D)[36],
                                   double& cM) {
  const int64_t pt = e * 8 + gp;
  double cK, c0, cc;
  operator_factors(p, a.k, cK, c0, cc, cM);
  double D0[36];
  initial_D(p, D0);
  double Dc[36];
  if (cc != 0.0) pimy_committed_tangent(p, a.C + pt, a.npts, Dc);
  int u = 0;
  double Dt[36];
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j, ++u) Dt[6 * i + j] = Dt[6 * j + i] = a.tangent[tangent_at(a, e, gp, u)];
  for (int i = 0; i < 36; ++i) {
    double v = cK * Dt[i];
    if (c0 != 0.0) v += c0 * D0[i];
    if (cc != 0.0) v += cc * Dc[i];
    D[i] = v;
  }
}

__device__ inline bool stiffness_damping(const PimyParams& p) {
  return p.betaK != 0.0 || p.betaKinit != 0.0 || p.betaKcomm != 0.0;
}

// N_j at Gauss point gp as gauss_shape forms it (the same operations on the same constants, so
// the same bits): the summing lane forms the body terms dvol b N_j from each point's dvol.
__device__ __forceinline__ double gauss_N(int gp, int k) {
  const double r3 = 0.57735026918962576451;
  const double xi = (gp & 4) ? r3 : -r3, eta = (gp & 2) ? r3 : -r3, zeta = (gp & 1) ? r3 : -r3;
  const double fx = kSx[k] > 0 ? 1.0 + xi : 1.0 - xi;
  const double fy = kSy[k] > 0 ? 1.0 + eta : 1.0 - eta;
  const double fz = kSz[k] > 0 ? 1.0 + zeta : 1.0 - zeta;
  const double cxy = 0.125 * fx * fy;
  return cxy * fz;
}

// The internal force: the update at u at every point (STORE also keeps the tangents; REPLAY takes
// each point's stress and tangent from the compacted pass instead of updating; DAMP adds the
// Rayleigh damping force at the velocities a.v, std_brick.hpp:409-464). One thread per point.
// A point leaves in LDS its node forces (and damping terms) and its dvol, 25 (49) values: the
// element's first lane forms the body terms itself. The registers are held to 256, two waves per
// SIMD (the return map's history loads then overlap another wave's arithmetic): the element's
// geometry is formed again after the return map instead of held through it, the tangent kept as
// its parts (coefficient, deviator, G, K) and the damping's D, D_0 and D_c formed entry by entry.
// DAMP blocks hold half the points.
template <int MODE, bool STORE, bool REPLAY = false, bool DAMP = false>
__global__ __launch_bounds__(kPointsPerBlock, 2) void pimy_batch_kernel(PimyBatchArgs a) {
  static_assert(MODE == 0, "the operator and its diagonal run in pimy_operator_kernel");
  constexpr int PTS = DAMP ? kPointsPerBlock / 2 : kPointsPerBlock, W = DAMP ? 6 : 3;
  // per point: 8 nodes x (3 force terms [, 3 damping terms]), then dvol
  __shared__ double part[PTS][8 * W + 1];
  const int t = threadIdx.x;  // the launch has PTS threads
  const int e = blockIdx.x * (PTS / 8) + t / 8, gp = t % 8;
  const bool live = e < a.count;
  if (live) {
    const PimyParams& p = a.params[a.param[e]];
    const int64_t pt = static_cast<int64_t>(e) * 8 + gp;
    double* mine = part[t];
    if constexpr (MODE == 0) {
      // the point's tangent as its parts (entries formed where they are used, the same values)
      double sig[6];
      PimyTangentParts Dt{};
      if constexpr (REPLAY) {
        for (int c = 0; c < 6; ++c) sig[c] = a.cres[c * a.npts + pt];
        if constexpr (STORE)
          for (int w = 0; w < 21; ++w) a.tangent[tangent_at(a, e, gp, w)] = a.cres[(6 + w) * a.npts + pt];
      } else {
        // the strain in a scope of its own: the geometry is formed again after the return map
        // rather than held through it (the same operations, so the same values)
        double eps[6];
        {
          double X[8], Y[8], Z[8], u[24];
          int eq[24];
          element_geometry(a, e, X, Y, Z, eq);
          Shape<double> s;
          gauss_shape(X, Y, Z, gp, s);
          for (int k = 0; k < 24; ++k) u[k] = eq[k] >= 0 ? a.x[eq[k]] : 0.0;
          point_strain(s, u, eps);
        }
        pimy_update_parts(p, a.C + pt, a.T + pt, a.npts, eps, sig, Dt, a.touched + pt);
        if constexpr (STORE) {
          int w = 0;
          for (int i = 0; i < 6; ++i)
            for (int j = i; j < 6; ++j, ++w) a.tangent[tangent_at(a, e, gp, w)] = Dt.entry(p.load_stage, i, j);
        }
      }
      double X[8], Y[8], Z[8];
      int eq[24];
      element_geometry(a, e, X, Y, Z, eq);
      Shape<double> s;
      gauss_shape(X, Y, Z, gp, s);
      for (int j = 0; j < 8; ++j) {
        double rj[3];
        node_force(s, sig, j, rj);
        for (int q = 0; q < 3; ++q) mine[W * j + q] = rj[q];
      }
      mine[8 * W] = s.dvol;  // the body terms s.dvol b N_j are formed by the summing lane
      if constexpr (DAMP) {
        // sigma_d = (betaK D + betaKinit D_0 + betaKcomm D_c) B v; the node terms B_j^T sigma_d dvol
        // + alphaM N_j dvol rho v (point_update_shape :420-442, point_node_contribution :448-464)
        const bool stiff = stiffness_damping(p);
        double vel[24], dsig[6] = {0, 0, 0, 0, 0, 0}, vgp[3] = {0, 0, 0};
        for (int k = 0; k < 24; ++k) vel[k] = eq[k] >= 0 ? a.v[eq[k]] : 0.0;
        if (stiff) {
          // D_0 and D_c entry by entry (the same values as the matrices, without holding them)
          double rate[6];
          PimyTangentParts Dc{};
          if (p.betaKcomm != 0.0) Dc = pimy_committed_tangent_parts(p, a.C + pt, a.npts);
          point_strain(s, vel, rate);
          for (int r = 0; r < 6; ++r) {
            double sum = 0.0;
            for (int c = 0; c < 6; ++c) {
              double Drc = p.betaK * Dt.entry(p.load_stage, r, c);
              if (p.betaKinit != 0.0) Drc += p.betaKinit * initial_entry(p, r, c);
              if (p.betaKcomm != 0.0) Drc += p.betaKcomm * Dc.entry(p.load_stage, r, c);
              sum += Drc * rate[c];
            }
            dsig[r] = sum;
          }
        }
        if (p.alphaM != 0.0 && p.mass_type == 0)
          for (int j = 0; j < 8; ++j)
            for (int q = 0; q < 3; ++q) vgp[q] += s.N[j] * vel[3 * j + q];
        for (int j = 0; j < 8; ++j) {
          double dj[3] = {0, 0, 0};
          if (stiff) node_force(s, dsig, j, dj);
          if (p.alphaM != 0.0) {
            const double temp = s.N[j] * s.dvol * p.rho;
            for (int q = 0; q < 3; ++q) dj[q] += p.alphaM * temp * (p.mass_type == 0 ? vgp[q] : vel[3 * j + q]);
          }
          for (int q = 0; q < 3; ++q) mine[W * j + 3 + q] = dj[q];
        }
      }
    }
  }
  __syncthreads();
  if (!live || gp != 0) return;
  double out[24], fdamp[24];
  for (int k = 0; k < 24; ++k) out[k] = fdamp[k] = 0.0;
  const PimyParams& p = a.params[a.param[e]];
  for (int g = 0; g < 8; ++g) {
    const double* c = part[t + g];
    const double dvol = c[8 * W];
    for (int j = 0; j < 8; ++j) {
      const double Nj = gauss_N(g, j);
      for (int q = 0; q < 3; ++q) {
        out[3 * j + q] += c[W * j + q];
        out[3 * j + q] -= dvol * p.b[q] * Nj;  // body
        if constexpr (DAMP) fdamp[3 * j + q] += c[W * j + 3 + q];
      }
    }
  }
  if constexpr (DAMP)  // fe = resid + fdamp where the element damps
    if (stiffness_damping(p) || p.alphaM != 0.0)
      for (int k = 0; k < 24; ++k) out[k] += fdamp[k];
  const size_t P = static_cast<size_t>(a.padded);
  for (int k = 0; k < 24; ++k) a.slots[k * P + e] = out[k];
}

// diag(B^T D_eff B dvol + cM M) at one point, tied DOFs of the element merged onto the first
// (element_operator_diagonal, element_operator.hpp:138-209)
__device__ inline void element_point_diagonal(const Shape<double>& s, const PimyParams& p, const double (&D)[36],
                                              double cM, const int (&eq)[24], double (&de)[24]) {
  for (int k = 0; k < 24; ++k) de[k] = 0.0;
  // B_J^T (D dvol) B_K entry (r, c): with g the gradients, the isotropic-free general form
  auto block = [&](int J, int K, int r, int c) {
    double BJ[6][3] = {}, BK[6][3] = {};
    const double gJ[3] = {s.dx[J], s.dy[J], s.dz[J]}, gK[3] = {s.dx[K], s.dy[K], s.dz[K]};
    BJ[0][0] = gJ[0]; BJ[1][1] = gJ[1]; BJ[2][2] = gJ[2]; BJ[3][0] = gJ[1]; BJ[3][1] = gJ[0];
    BJ[4][1] = gJ[2]; BJ[4][2] = gJ[1]; BJ[5][0] = gJ[2]; BJ[5][2] = gJ[0];
    BK[0][0] = gK[0]; BK[1][1] = gK[1]; BK[2][2] = gK[2]; BK[3][0] = gK[1]; BK[3][1] = gK[0];
    BK[4][1] = gK[2]; BK[4][2] = gK[1]; BK[5][0] = gK[2]; BK[5][2] = gK[0];
    double o = 0.0;
    for (int m = 0; m < 6; ++m) {
      double bd = 0.0;
      for (int l = 0; l < 6; ++l) bd += BJ[l][r] * (D[6 * l + m] * s.dvol);
      o += bd * BK[m][c];
    }
    return o;
  };
  auto mass_entry = [&](int J, int K) {
    const double temp = s.N[J] * s.dvol * p.rho;
    if (p.mass_type == 0) return temp * s.N[K];
    return J == K ? temp : 0.0;
  };
  int first[24];
  bool merged = false;
  for (int k = 0; k < 24; ++k) {
    first[k] = k;
    if (eq[k] < 0) continue;
    for (int b = 0; b < k; ++b)
      if (eq[b] == eq[k]) {
        first[k] = first[b];
        merged = true;
        break;
      }
  }
  for (int j = 0; j < 8; ++j) {
    const double m = cM != 0.0 ? cM * mass_entry(j, j) : 0.0;
    for (int q = 0; q < 3; ++q) de[3 * j + q] += block(j, j, q, q) + m;
  }
  if (merged)
    for (int J = 0; J < 8; ++J)
      for (int K = 0; K < 8; ++K)
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            const int ia = 3 * J + r, ic = 3 * K + c;
            if (ia == ic || eq[ia] < 0 || eq[ia] != eq[ic]) continue;
            const double m = (cM != 0.0 && J != K && r == c) ? cM * mass_entry(J, K) : 0.0;
            de[first[ia]] += block(J, K, r, c) + m;
          }
}

// The operator (DIAG false) and its diagonal on the stored tangents: one thread per element, its
// 8 points in order, each point's node terms added to the element's 24 values as they come
// (element_operator_apply / _diagonal, element_operator.hpp:82-209). Registers only. COMM: some
// parameter set has a committed-tangent term (betaKcomm with damping in the coefficients).
template <bool DIAG, bool COMM>
__global__ __launch_bounds__(128) void pimy_operator_kernel(PimyBatchArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.count) return;
  double X[8], Y[8], Z[8];
  int eq[24];
  element_geometry(a, e, X, Y, Z, eq);
  const PimyParams& p = a.params[a.param[e]];
  double v[24];
  if constexpr (!DIAG)
    for (int k = 0; k < 24; ++k) v[k] = operator_input(a, eq, k);
  double out[24];
  for (int k = 0; k < 24; ++k) out[k] = 0.0;
  double cK, c0, cc, cM;
  operator_factors(p, a.k, cK, c0, cc, cM);
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    Shape<double> s;
    gauss_shape(X, Y, Z, gp, s);
    if constexpr (!DIAG) {
      // sigma = D_eff eps with D_eff's entries formed one at a time as effective_D forms them
      double eps[6], sig[6], Dc[36];
      if constexpr (COMM)
        if (cc != 0.0) pimy_committed_tangent(p, a.C + static_cast<int64_t>(e) * 8 + gp, a.npts, Dc);
      point_strain(s, v, eps);
      for (int r = 0; r < 6; ++r) {
        double sum = 0.0;
        for (int c = 0; c < 6; ++c) {
          const int lo = r < c ? r : c, hi = r < c ? c : r;
          const int w = lo * 6 - lo * (lo - 1) / 2 + (hi - lo);  // packed upper triangle, row-major
          double d = cK * a.tangent[tangent_at(a, e, gp, w)];
          if (c0 != 0.0) d += c0 * initial_entry(p, r, c);
          if constexpr (COMM)
            if (cc != 0.0) d += cc * Dc[6 * r + c];
          sum += d * eps[c];
        }
        sig[r] = sum;
      }
      double g[3] = {0, 0, 0};
      if (cM != 0.0 && p.mass_type == 0)
        for (int kk = 0; kk < 8; ++kk)
          for (int q = 0; q < 3; ++q) g[q] += s.N[kk] * v[3 * kk + q];
      for (int j = 0; j < 8; ++j) {
        double fj[3];
        node_force(s, sig, j, fj);
        const double temp = s.N[j] * s.dvol * p.rho;
        for (int q = 0; q < 3; ++q) {
          out[3 * j + q] += fj[q];
          if (cM != 0.0) out[3 * j + q] += cM * (temp * (p.mass_type == 0 ? g[q] : v[3 * j + q]));
        }
      }
    } else {
      double D[36], cMd;
      effective_D(a, p, e, gp, D, cMd);
      double de[24];
      element_point_diagonal(s, p, D, cMd, eq, de);
      for (int k = 0; k < 24; ++k) out[k] += de[k];
    }
  }
  const size_t P = static_cast<size_t>(a.padded);
  for (int k = 0; k < 24; ++k) a.slots[k * P + e] = out[k];
}

// The same for small batches, one thread per point (8 lanes per element), so that a batch of a
// few hundred elements still runs thousands of threads: each point leaves its node terms in
// LDS and the element's first lane adds them in point order (the same sums as above).
template <bool DIAG>
__global__ __launch_bounds__(kPointsPerBlock) void pimy_point_operator_kernel(PimyBatchArgs a) {
  __shared__ double part[kPointsPerBlock][48];  // per point: 8 nodes x (3 stiffness, 3 mass) or 24 x 2
  const int t = threadIdx.x;
  const int e = blockIdx.x * kElemsPerBlock + t / 8, gp = t % 8;
  const bool live = e < a.count;
  if (live) {
    double X[8], Y[8], Z[8];
    int eq[24];
    element_geometry(a, e, X, Y, Z, eq);
    const PimyParams& p = a.params[a.param[e]];
    Shape<double> s;
    gauss_shape(X, Y, Z, gp, s);
    double* mine = part[t];
    double D[36], cM;
    effective_D(a, p, e, gp, D, cM);
    if constexpr (!DIAG) {
      double v[24], eps[6], sig[6];
      for (int k = 0; k < 24; ++k) v[k] = operator_input(a, eq, k);
      point_strain(s, v, eps);
      for (int r = 0; r < 6; ++r) {
        double sum = 0.0;
        for (int c = 0; c < 6; ++c) sum += D[6 * r + c] * eps[c];
        sig[r] = sum;
      }
      double g[3] = {0, 0, 0};
      if (cM != 0.0 && p.mass_type == 0)
        for (int kk = 0; kk < 8; ++kk)
          for (int q = 0; q < 3; ++q) g[q] += s.N[kk] * v[3 * kk + q];
      for (int j = 0; j < 8; ++j) {
        double fj[3];
        node_force(s, sig, j, fj);
        const double temp = s.N[j] * s.dvol * p.rho;
        for (int q = 0; q < 3; ++q) {
          mine[6 * j + q] = fj[q];
          mine[6 * j + 3 + q] = cM != 0.0 ? cM * (temp * (p.mass_type == 0 ? g[q] : v[3 * j + q])) : 0.0;
        }
      }
    } else {
      double de[24];
      element_point_diagonal(s, p, D, cM, eq, de);
      for (int k = 0; k < 24; ++k) {
        mine[2 * k] = de[k];
        mine[2 * k + 1] = 0.0;
      }
    }
  }
  __syncthreads();
  if (!live || gp != 0) return;
  const PimyParams& p = a.params[a.param[e]];
  double cK, c0, cc, cM;
  operator_factors(p, a.k, cK, c0, cc, cM);
  double out[24];
  for (int k = 0; k < 24; ++k) out[k] = 0.0;
  for (int g = 0; g < 8; ++g) {
    const double* c = part[t + g];
    if constexpr (!DIAG) {
      for (int j = 0; j < 8; ++j)
        for (int q = 0; q < 3; ++q) {
          out[3 * j + q] += c[6 * j + q];
          if (cM != 0.0) out[3 * j + q] += c[6 * j + 3 + q];
        }
    } else {
      for (int k = 0; k < 24; ++k) out[k] += c[2 * k];
    }
  }
  const size_t P = static_cast<size_t>(a.padded);
  for (int k = 0; k < 24; ++k) a.slots[k * P + e] = out[k];
}

// The compacted return map (compacted_kernels.hpp:85-202). Predict: one thread per Gauss point,
// the strain as the fused kernel forms it, the elastic predictor; an elastic point is finished
// (stress and tangent stored), a plastic one keeps its strain and joins the list (one atomic per
// wave). Correct: a fixed grid walks the list, one full update per listed point (the same
// function as the fused kernel, so the same bits). The force kernel then replays the stored
// results.
__global__ __launch_bounds__(kPointsPerBlock) void pimy_predict_kernel(PimyBatchArgs a) {
  const int t = threadIdx.x;
  const int e = blockIdx.x * kElemsPerBlock + t / 8, gp = t % 8;
  bool plastic = false;
  const int64_t pt = static_cast<int64_t>(e) * 8 + gp;
  if (e < a.count) {
    double X[8], Y[8], Z[8];
    int eq[24];
    element_geometry(a, e, X, Y, Z, eq);
    const PimyParams& p = a.params[a.param[e]];
    Shape<double> s;
    gauss_shape(X, Y, Z, gp, s);
    double u[24], eps[6], sig[6], D[36];
    for (int k = 0; k < 24; ++k) u[k] = eq[k] >= 0 ? a.x[eq[k]] : 0.0;
    point_strain(s, u, eps);
    plastic = pimy_predict(p, a.C + pt, a.T + pt, a.npts, eps, sig, D, a.touched + pt);
    if (plastic) {
      for (int c = 0; c < 6; ++c) a.cres[(27 + c) * a.npts + pt] = eps[c];
    } else {
      for (int c = 0; c < 6; ++c) a.cres[c * a.npts + pt] = sig[c];
      int w = 0;
      for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j, ++w) a.cres[(6 + w) * a.npts + pt] = D[6 * i + j];
    }
  }
  const uint64_t mask = __ballot(plastic);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(a.ccount, __popcll(mask));
  base = __shfl(base, leader, 64);
  if (plastic) a.clist[base + __popcll(mask & ((uint64_t(1) << lane) - 1))] = static_cast<int32_t>(pt);
}

__global__ __launch_bounds__(128) void pimy_correct_kernel(PimyBatchArgs a) {
  const int32_t n = *a.ccount;
  for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int64_t pt = a.clist[k];
    const PimyParams& p = a.params[a.param[pt / 8]];
    double eps[6], sig[6], D[36];
    for (int c = 0; c < 6; ++c) eps[c] = a.cres[(27 + c) * a.npts + pt];
    pimy_update(p, a.C + pt, a.T + pt, a.npts, eps, sig, D, a.touched + pt);
    for (int c = 0; c < 6; ++c) a.cres[c * a.npts + pt] = sig[c];
    int w = 0;
    for (int i = 0; i < 6; ++i)
      for (int j = i; j < 6; ++j, ++w) a.cres[(6 + w) * a.npts + pt] = D[6 * i + j];
  }
}

__global__ void pimy_clear_count_kernel(int32_t* count) {
  if (threadIdx.x == 0) count[0] = 0;
}

// y[q] = the sum of equation q's element slots, in (element, DOF) order (assembly.cpp:64-95).
// An equation with more than kLongGather slots (a tied level of a soil column) is left to
// long_gather_kernel.
constexpr int kLongGather = 64;

template <typename T>
__global__ void slot_gather_kernel(int32_t neq, const int32_t* __restrict__ start, const int32_t* __restrict__ list,
                                   const T* __restrict__ slots, T* __restrict__ y, const int32_t* done,
                                   int accumulate) {
  if (done && *done) return;
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= neq) return;
  const int32_t j0 = start[q], j1 = start[q + 1];
  if (j1 - j0 > kLongGather) return;
  T s = T(0);
  for (int32_t j = j0; j < j1; ++j) s += slots[list[j]];
  y[q] = accumulate ? y[q] + s : s;
}

// The long equations: one wave each, lane l summing slots l, l + 64, ... in order, then a fixed
// butterfly over the lanes.
template <typename T>
__global__ __launch_bounds__(256) void long_gather_kernel(int32_t nlong, const int32_t* __restrict__ eqs,
                                                          const int32_t* __restrict__ start,
                                                          const int32_t* __restrict__ list,
                                                          const T* __restrict__ slots, T* __restrict__ y,
                                                          const int32_t* done, int accumulate) {
  const int u = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x) / 64, lane = threadIdx.x % 64;
  if (u >= nlong) return;  // the whole wave
  if (done && *done) return;
  const int32_t q = eqs[u];
  T s = T(0);
  for (int32_t j = start[q] + lane; j < start[q + 1]; j += 64) s += slots[list[j]];
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[q] = accumulate ? y[q] + s : s;
}

__global__ void pimy_batch_init_kernel(int64_t npts, const PimyParams* __restrict__ params,
                                       const int32_t* __restrict__ param, double* S) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const PimyParams& p = params[param[i / 8]];
  for (int k = 0; k < kPimyNState; ++k) S[k * npts + i] = 0.0;
  S[kPG * npts + i] = p.G;
  S[kPK * npts + i] = p.K;
  for (int j = 1; j <= kPimyNS; ++j) {
    S[(kPSizes + j - 1) * npts + i] = p.size[j];
    S[(kPModuli + j - 1) * npts + i] = p.modulus[j];
  }
}

PimyBatchArgs batch_args(fx_batch* b, const double* x, const fx_coeffs& k, const double* v = nullptr) {
  PimyBatchArgs a{};
  a.count = b->count;
  a.padded = b->padded;
  a.nnode = b->nnode;
  a.npts = b->npts;
  a.node = b->node;
  a.param = b->param;
  a.node_eq = b->node_eq;
  a.coords = b->coords;
  a.params = b->pimy_params;
  a.C = b->state_c;
  a.T = b->state_t;
  a.tangent = b->tangent;
  a.x = x;
  a.v = v;
  a.touched = b->pimy_touched;
  a.slots = b->slots;
  a.k = k;
  a.cres = b->cres;
  a.clist = b->clist;
  a.ccount = b->ccount;
  a.unit_dir = -1;
  a.ssp = b->ssp_elem;
  a.ssp_stride = static_cast<int64_t>(b->blk.nblocks) * kBlockElems;
  a.ssp_pos = b->blk.pos;
  return a;
}

// y[q] = the sum of equation q's element slots (slot_gather_kernel, long_gather_kernel)
template <typename T>
int gather(fx_batch* b, const T* slots, T* y, const int32_t* done, int accumulate = 0) {
  if (b->neq > 0)
    hipLaunchKernelGGL(slot_gather_kernel<T>, dim3((b->neq + 255) / 256), dim3(256), 0, b->ctx->stream, b->neq,
                       b->gather_start, b->gather_list, slots, y, done, accumulate);
  if (b->nlong > 0)
    hipLaunchKernelGGL(long_gather_kernel<T>, dim3((b->nlong + 3) / 4), dim3(256), 0, b->ctx->stream, b->nlong,
                       b->long_eqs, b->gather_start, b->gather_list, slots, y, done, accumulate);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

// The FP32 copy of the stored tangents that the FP32 operator reads (half the bytes).
__global__ void tangent_to_f32_kernel(int64_t n, const double* __restrict__ t, float* __restrict__ t32) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) t32[i] = static_cast<float>(t[i]);
}

// The FP32 operator on the stored tangents, for the inner solves of mixed-precision CG: one
// thread per element as pimy_operator_kernel, its terms in the same order, with x, the
// tangents, strains, stresses and sums in FP32. The geometry (Jacobian, shape derivatives,
// dvol), the operator factors and the committed tangent are formed in FP64 and rounded once.
template <bool COMM>
__global__ __launch_bounds__(128) void pimy_operator32_kernel(PimyBatchArgs a, const float* __restrict__ x,
                                                              const float* __restrict__ tangent, float* slots) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.count) return;
  double X[8], Y[8], Z[8];
  int eq[24];
  element_geometry(a, e, X, Y, Z, eq);
  const PimyParams& p = a.params[a.param[e]];
  float v[24], out[24];
  for (int k = 0; k < 24; ++k) {
    v[k] = eq[k] >= 0 ? x[eq[k]] : 0.0f;
    out[k] = 0.0f;
  }
  double cK, c0, cc, cM;
  operator_factors(p, a.k, cK, c0, cc, cM);
  const float fK = static_cast<float>(cK), fM = static_cast<float>(cM);
  float D0[36];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) D0[6 * r + c] = static_cast<float>(c0 * initial_entry(p, r, c));
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    Shape<double> s;
    gauss_shape(X, Y, Z, gp, s);
    float dx[8], dy[8], dz[8];
    for (int j = 0; j < 8; ++j) {
      dx[j] = static_cast<float>(s.dx[j]);
      dy[j] = static_cast<float>(s.dy[j]);
      dz[j] = static_cast<float>(s.dz[j]);
    }
    const float dvol = static_cast<float>(s.dvol);
    float eps[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 8; ++j) {
      const float u0 = v[3 * j], u1 = v[3 * j + 1], u2 = v[3 * j + 2];
      eps[0] += dx[j] * u0;
      eps[1] += dy[j] * u1;
      eps[2] += dz[j] * u2;
      eps[3] += dy[j] * u0 + dx[j] * u1;
      eps[4] += dz[j] * u1 + dy[j] * u2;
      eps[5] += dz[j] * u0 + dx[j] * u2;
    }
    double Dc[36];
    if constexpr (COMM)
      if (cc != 0.0) pimy_committed_tangent(p, a.C + static_cast<int64_t>(e) * 8 + gp, a.npts, Dc);
    float sig[6];
    for (int r = 0; r < 6; ++r) {
      float sum = 0.0f;
      for (int c = 0; c < 6; ++c) {
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        const int w = lo * 6 - lo * (lo - 1) / 2 + (hi - lo);  // packed upper triangle, row-major
        float d = fK * tangent[tangent_at(a, e, gp, w)] + D0[6 * r + c];
        if constexpr (COMM)
          if (cc != 0.0) d += static_cast<float>(cc * Dc[6 * r + c]);
        sum += d * eps[c];
      }
      sig[r] = sum * dvol;
    }
    float g[3] = {0, 0, 0};
    if (cM != 0.0 && p.mass_type == 0)
      for (int kk = 0; kk < 8; ++kk)
        for (int q = 0; q < 3; ++q) g[q] += static_cast<float>(s.N[kk]) * v[3 * kk + q];
    for (int j = 0; j < 8; ++j) {
      out[3 * j] += dx[j] * sig[0] + dy[j] * sig[3] + dz[j] * sig[5];
      out[3 * j + 1] += dy[j] * sig[1] + dx[j] * sig[3] + dz[j] * sig[4];
      out[3 * j + 2] += dz[j] * sig[2] + dy[j] * sig[4] + dx[j] * sig[5];
      if (cM != 0.0) {
        const float temp = static_cast<float>(s.N[j] * s.dvol * p.rho);
        for (int q = 0; q < 3; ++q) out[3 * j + q] += fM * (temp * (p.mass_type == 0 ? g[q] : v[3 * j + q]));
      }
    }
  }
  const size_t P = static_cast<size_t>(a.padded);
  for (int k = 0; k < 24; ++k) slots[k * P + e] = out[k];
}

// ---- SSPbrick<PressureIndependMultiYield>: one material point per element (the element's point 0
// of the batch's eight-point history), one thread per element (ssp_brick.hpp:1056-1156) ----------
struct SspRef {
  const double* c;
  int64_t stride, at;
  __device__ double operator[](int k) const { return c[k * stride + at]; }
};

__device__ inline SspRef ssp_ref(const PimyBatchArgs& a, int e) {
  return SspRef{a.ssp, a.ssp_stride, static_cast<int64_t>(a.ssp_pos[e])};
}

// eps = Bnot u, engineering shear (point_strain, ssp_brick.hpp:1018-1033)
__device__ inline void ssp_mean_strain(const SspRef& c, const double (&u)[24], double (&eps)[6]) {
  for (int k = 0; k < 6; ++k) eps[k] = 0.0;
  for (int i = 0; i < 8; ++i) {
    const double bx = c[3 * i], by = c[3 * i + 1], bz = c[3 * i + 2];
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

// mVol Bnot_j^T sigma (point_node_force, :1035-1041)
__device__ inline void ssp_node_force(const SspRef& c, const double (&sig)[6], int j, double (&f)[3]) {
  const double bx = c[3 * j], by = c[3 * j + 1], bz = c[3 * j + 2], V = c[24];
  f[0] = (bx * sig[0] + by * sig[3] + bz * sig[5]) * V;
  f[1] = (by * sig[1] + bx * sig[3] + bz * sig[4]) * V;
  f[2] = (bz * sig[2] + by * sig[4] + bx * sig[5]) * V;
}

// interior(r, q) from its packed upper triangle
__device__ inline double ssp_int(const SspRef& c, int r, int q) {
  if (r > q) {
    const int t = r;
    r = q;
    q = t;
  }
  return c[kSspInterior + 12 * r - r * (r - 1) / 2 + (q - r)];
}

// Kstab v at the 24 DOFs: z = interior (Mben v), node j gets Mben_j^T z (element_term_reduce /
// _node, :964-989)
__device__ inline void ssp_kstab(const SspRef& c, const double (&v)[24], double (&out)[24]) {
  double w[12], z[12];
  for (int m = 0; m < 4; ++m)
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int i = 0; i < 8; ++i) s += c[kSspGamma + 8 * m + i] * v[3 * i + d];
      w[3 * m + d] = s;
    }
  for (int r = 0; r < 12; ++r) {
    double s = 0.0;
    for (int q = 0; q < 12; ++q) s += ssp_int(c, r, q) * w[q];
    z[r] = s;
  }
  for (int j = 0; j < 8; ++j)
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int m = 0; m < 4; ++m) s += c[kSspGamma + 8 * m + j] * z[3 * m + d];
      out[3 * j + d] = s;
    }
}

// Kstab(a, b), in the order of (min, max) (element_term_entry, :991-1005)
__device__ inline double ssp_kstab_entry(const SspRef& c, int a, int b) {
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  const int i = a / 3, d = a % 3, k = b / 3, f = b % 3;
  double s = 0.0;
  for (int g = 0; g < 4; ++g)
    for (int h = 0; h < 4; ++h) s += c[kSspGamma + 8 * g + i] * ssp_int(c, 3 * g + d, 3 * h + f) * c[kSspGamma + 8 * h + k];
  return s;
}

// (mVol Bnot_J^T D Bnot_K)[p][q] (point_block_tangent, :1049-1067)
__device__ inline double ssp_block_entry(const SspRef& c, const double (&D)[36], int J, int K, int p, int q) {
  auto B = [&](int n, int r, int col) {  // Bnot_n[r][col]
    const double bx = c[3 * n], by = c[3 * n + 1], bz = c[3 * n + 2];
    switch (r) {
      case 0: return col == 0 ? bx : 0.0;
      case 1: return col == 1 ? by : 0.0;
      case 2: return col == 2 ? bz : 0.0;
      case 3: return col == 0 ? by : col == 1 ? bx : 0.0;
      case 4: return col == 1 ? bz : col == 2 ? by : 0.0;
      default: return col == 0 ? bz : col == 2 ? bx : 0.0;
    }
  };
  const double V = c[24];
  double s = 0.0;
  for (int r = 0; r < 6; ++r) {
    double bd = 0.0;
    for (int t = 0; t < 6; ++t) bd += B(J, t, p) * (D[6 * t + r] * V);
    s += bd * B(K, r, q);
  }
  return s;
}

// The force (STORE: also the point's tangent): Kstab u, then mVol Bnot^T sigma - the body force,
// and with velocities the damping force (betaK + betaKinit + betaKcomm) Kstab v + mVol Bnot^T
// sigma_d + alphaM m v, sigma_d = (betaK D + betaKinit D + betaKcomm D_c) Bnot v (the initial
// stiffness is the tangent); fe = resid + fdamp (compute_with, :1107-1156).
template <bool STORE>
__global__ __launch_bounds__(128) void pimy_ssp_kernel(PimyBatchArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.count) return;
  const size_t P = static_cast<size_t>(a.padded);
  int eq[24];
  for (int k = 0; k < 8; ++k) {
    const int n = a.node[k * P + e];
    for (int d = 0; d < 3; ++d) eq[3 * k + d] = a.node_eq[3 * n + d];
  }
  const PimyParams& p = a.params[a.param[e]];
  const SspRef c = ssp_ref(a, e);
  const bool stiff = stiffness_damping(p);
  const bool damped = a.v && (stiff || p.alphaM != 0.0);
  double u[24], v[24], resid[24], fdamp[24];
  for (int k = 0; k < 24; ++k) {
    u[k] = eq[k] >= 0 ? a.x[eq[k]] : 0.0;
    v[k] = (damped && eq[k] >= 0) ? a.v[eq[k]] : 0.0;
  }
  ssp_kstab(c, u, resid);
  const double beta = p.betaK + p.betaKinit + p.betaKcomm;
  if (damped && beta != 0.0) {
    ssp_kstab(c, v, fdamp);
    for (int k = 0; k < 24; ++k) fdamp[k] = beta * fdamp[k];
  } else {
    for (int k = 0; k < 24; ++k) fdamp[k] = 0.0;
  }
  const int64_t pt = static_cast<int64_t>(e) * 8;
  double eps[6], sig[6], D[36], dsig[6] = {0, 0, 0, 0, 0, 0};
  ssp_mean_strain(c, u, eps);
  pimy_update(p, a.C + pt, a.T + pt, a.npts, eps, sig, D, a.touched + pt);
  if constexpr (STORE) {
    int w = 0;
    for (int i = 0; i < 6; ++i)
      for (int j = i; j < 6; ++j, ++w) a.tangent[tangent_at(a, e, 0, w)] = D[6 * i + j];
  }
  if (damped && stiff) {
    double Dc[36], rate[6];
    if (p.betaKcomm != 0.0) pimy_committed_tangent(p, a.C + pt, a.npts, Dc);
    ssp_mean_strain(c, v, rate);
    for (int i = 0; i < 6; ++i) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) {
        double Diq = p.betaK * D[6 * i + q];
        if (p.betaKinit != 0.0) Diq += p.betaKinit * D[6 * i + q];
        if (p.betaKcomm != 0.0) Diq += p.betaKcomm * Dc[6 * i + q];
        s += Diq * rate[q];
      }
      dsig[i] = s;
    }
  }
  double out[24];
  for (int j = 0; j < 8; ++j) {
    double fj[3], dj[3] = {0, 0, 0};
    ssp_node_force(c, sig, j, fj);
    const double wj = c[25 + j];
    if (damped) {
      if (stiff) ssp_node_force(c, dsig, j, dj);
      if (p.alphaM != 0.0) {
        const double m = p.rho * wj;
        for (int q = 0; q < 3; ++q) dj[q] += p.alphaM * m * v[3 * j + q];
      }
    }
    for (int q = 0; q < 3; ++q) {
      double r = resid[3 * j + q];
      r += fj[q];
      r -= p.b[q] * wj;
      const double fd = fdamp[3 * j + q] + dj[q];
      out[3 * j + q] = damped ? r + fd : r;
    }
  }
  for (int k = 0; k < 24; ++k) a.slots[k * P + e] = out[k];
}

// The operator on the stored tangent (DIAG: its diagonal, merged DOFs' cross terms on the first):
// mVol Bnot^T D_eff Bnot + cM m + cE Kstab, D_eff = cK D_t + cc D_c with cK = cS + cC (betaK +
// betaKinit), cE = cS + cC (betaK + betaKinit + betaKcomm) (element_operator.hpp:45-56 with the
// initial stiffness taken as the tangent; element_term.hpp:34-37).
template <bool DIAG>
__global__ __launch_bounds__(128) void pimy_ssp_operator_kernel(PimyBatchArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.count) return;
  const size_t P = static_cast<size_t>(a.padded);
  int eq[24];
  for (int k = 0; k < 8; ++k) {
    const int n = a.node[k * P + e];
    for (int d = 0; d < 3; ++d) eq[3 * k + d] = a.node_eq[3 * n + d];
  }
  const PimyParams& p = a.params[a.param[e]];
  const SspRef c = ssp_ref(a, e);
  const fx_coeffs& k = a.k;
  const double cK = k.stiffness + k.damping * p.betaK + k.damping * p.betaKinit;
  const double cc = k.damping * p.betaKcomm;
  const double cM = k.mass + k.damping * p.alphaM;
  const double cE = k.stiffness + k.damping * (p.betaK + p.betaKinit + p.betaKcomm);
  const int64_t pt = static_cast<int64_t>(e) * 8;
  double Deff[36];
  {
    double Dc[36];
    if (cc != 0.0) pimy_committed_tangent(p, a.C + pt, a.npts, Dc);
    int w = 0;
    double Dt[36];
    for (int i = 0; i < 6; ++i)
      for (int j = i; j < 6; ++j, ++w) Dt[6 * i + j] = Dt[6 * j + i] = cK != 0.0 ? a.tangent[tangent_at(a, e, 0, w)] : 0.0;
    for (int i = 0; i < 36; ++i) {
      double v = cK * Dt[i];
      if (cc != 0.0) v += cc * Dc[i];
      Deff[i] = v;
    }
  }
  double out[24];
  if constexpr (DIAG) {
    int first[24];
    bool merged = false;
    for (int i = 0; i < 24; ++i) {
      out[i] = 0.0;
      first[i] = i;
      if (eq[i] < 0) continue;
      for (int j = 0; j < i; ++j)
        if (eq[j] == eq[i]) {
          first[i] = first[j];
          merged = true;
          break;
        }
    }
    for (int j = 0; j < 8; ++j) {
      const double m = cM != 0.0 ? cM * (p.rho * c[25 + j]) : 0.0;
      for (int q = 0; q < 3; ++q) out[3 * j + q] += ssp_block_entry(c, Deff, j, j, q, q) + m;
    }
    if (merged)
      for (int ia = 0; ia < 24; ++ia)
        for (int ib = 0; ib < 24; ++ib) {
          if (ia == ib || eq[ia] < 0 || eq[ia] != eq[ib]) continue;
          out[first[ia]] += ssp_block_entry(c, Deff, ia / 3, ib / 3, ia % 3, ib % 3);
        }
    if (cE != 0.0) {
      for (int i = 0; i < 24; ++i) out[i] += cE * ssp_kstab_entry(c, i, i);
      if (merged)
        for (int ia = 0; ia < 24; ++ia)
          for (int ib = 0; ib < 24; ++ib) {
            if (ia == ib || eq[ia] < 0 || eq[ia] != eq[ib]) continue;
            out[first[ia]] += cE * ssp_kstab_entry(c, ia, ib);
          }
    }
  } else {
    double v[24];
    for (int i = 0; i < 24; ++i) v[i] = operator_input(a, eq, i);
    double eps[6], sig[6];
    ssp_mean_strain(c, v, eps);
    for (int i = 0; i < 6; ++i) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) s += Deff[6 * i + q] * eps[q];
      sig[i] = s;
    }
    for (int j = 0; j < 8; ++j) {
      double fj[3];
      ssp_node_force(c, sig, j, fj);
      for (int q = 0; q < 3; ++q) {
        out[3 * j + q] = fj[q];
        if (cM != 0.0) out[3 * j + q] += cM * (p.rho * c[25 + j] * v[3 * j + q]);
      }
    }
    if (cE != 0.0) {
      double kv[24];
      ssp_kstab(c, v, kv);
      for (int i = 0; i < 24; ++i) out[i] += cE * kv[i];
    }
  }
  for (int i = 0; i < 24; ++i) a.slots[i * P + e] = out[i];
}

// The operator (DIAG false) or its diagonal on the stored tangents into a.slots, element values
// a * padded + e: one thread per point below kPointOperatorBelow elements, one per element above.
template <bool DIAG>
void launch_operator_slots(fx_batch* b, const PimyBatchArgs& a) {
  hipStream_t st = b->ctx->stream;
  if (b->ssp) {
    if (b->count > 0)
      hipLaunchKernelGGL(pimy_ssp_operator_kernel<DIAG>, dim3((b->count + 127) / 128), dim3(128), 0, st, a);
  } else if (b->count > 0 && b->count < kPointOperatorBelow) {  // small: one thread per point
    const dim3 grid((b->count + kElemsPerBlock - 1) / kElemsPerBlock);
    hipLaunchKernelGGL((pimy_point_operator_kernel<DIAG>), grid, dim3(kPointsPerBlock), 0, st, a);
  } else if (b->count > 0) {  // one element per thread
    bool comm = false;
    for (const PimyParams& q : b->host_pimy) comm |= a.k.damping * q.betaKcomm != 0.0;
    const dim3 g((b->count + 127) / 128);
    if (comm)
      hipLaunchKernelGGL((pimy_operator_kernel<DIAG, true>), g, dim3(128), 0, st, a);
    else
      hipLaunchKernelGGL((pimy_operator_kernel<DIAG, false>), g, dim3(128), 0, st, a);
  }
}

// MODE 0: the force (STORE: also the tangents); 1: the operator on the stored tangents; 2: its
// diagonal. The elements' 24 values go to their slots, then each equation gathers its slots.
template <int MODE, bool STORE>
int run_pimy(fx_batch* b, const PimyBatchArgs& a, double* y, const int32_t* done) {
  const dim3 grid((b->count + kElemsPerBlock - 1) / kElemsPerBlock);
  hipStream_t st = b->ctx->stream;
  if constexpr (MODE == 0) {
    if (b->ssp) {  // SSPbrick: one thread per element, its one point
      if (b->count > 0)
        hipLaunchKernelGGL(pimy_ssp_kernel<STORE>, dim3((b->count + 127) / 128), dim3(128), 0, st, a);
    } else if (b->count > 0 && a.v) {  // the damping force: the fused return map, half blocks
      const dim3 g2((b->count + kElemsPerBlock / 2 - 1) / (kElemsPerBlock / 2));
      hipLaunchKernelGGL((pimy_batch_kernel<0, STORE, false, true>), g2, dim3(kPointsPerBlock / 2), 0, st, a);
    } else if (b->count > 0 && b->cres) {  // compacted: predict, correct, replay
      hipLaunchKernelGGL(pimy_predict_kernel, grid, dim3(kPointsPerBlock), 0, st, a);
      const unsigned cgrid = static_cast<unsigned>(std::min<int64_t>((b->npts + 127) / 128, 2048));
      hipLaunchKernelGGL(pimy_correct_kernel, dim3(cgrid), dim3(128), 0, st, a);
      hipLaunchKernelGGL(pimy_clear_count_kernel, dim3(1), dim3(64), 0, st, a.ccount);
      hipLaunchKernelGGL((pimy_batch_kernel<0, STORE, true>), grid, dim3(kPointsPerBlock), 0, st, a);
    } else if (b->count > 0) {
      hipLaunchKernelGGL((pimy_batch_kernel<0, STORE>), grid, dim3(kPointsPerBlock), 0, st, a);
    }
  } else {
    launch_operator_slots<MODE == 2>(b, a);
  }
  return gather<double>(b, b->slots, y, done);
}

}  // namespace

int pimy_batch_init(fx_batch* b) {
  for (double* S : {b->state_c, b->state_t})
    if (b->npts > 0)
      hipLaunchKernelGGL(pimy_batch_init_kernel, dim3(static_cast<unsigned>((b->npts + 255) / 256)), dim3(256), 0,
                         b->ctx->stream, b->npts, b->pimy_params, b->param, S);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

int pimy_force64(fx_batch* b, const double* u, const double* v, double* f, bool store) {
  // the damping force only where some parameter set damps and velocities are given
  const PimyBatchArgs a = batch_args(b, u, fx_coeffs{1.0, 0.0, 0.0, 0}, b->pimy_damped ? v : nullptr);
  if (store) {
    if (int rc = run_pimy<0, true>(b, a, f, nullptr)) return rc;
    b->tangents_stored = true;
    if (b->tangent32 && b->npts > 0) {
      const int64_t nt = b->npts * 21;
      hipLaunchKernelGGL(tangent_to_f32_kernel, dim3(static_cast<unsigned>((nt + 255) / 256)), dim3(256), 0,
                         b->ctx->stream, nt, b->tangent, b->tangent32);
      FX_HIP(b->ctx, hipGetLastError());
    }
    return FX_OK;
  }
  return run_pimy<0, false>(b, a, f, nullptr);
}

int pimy_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y, const int32_t* done) {
  if (!b->tangents_stored && !k.initial && k.stiffness != 0.0) {
    b->ctx->error = "operator: no stored tangents (call fx_store_tangents first)";
    return FX_EINVAL;
  }
  return run_pimy<1, false>(b, batch_args(b, x, k), y, done);
}

int pimy_apply32(fx_batch* b, const fx_coeffs& k, const float* x, float* y, const int32_t* done) {
  if (!b->tangent32) {
    b->ctx->error = "operator: no FP32 operator (PressureIndependMultiYield batch created without FX_BATCH_FP32)";
    return FX_EUNSUP;
  }
  if (!b->tangents_stored && !k.initial && k.stiffness != 0.0) {
    b->ctx->error = "operator: no stored tangents (call fx_store_tangents first)";
    return FX_EINVAL;
  }
  const PimyBatchArgs a = batch_args(b, nullptr, k);
  float* slots = reinterpret_cast<float*>(b->slots);  // [24 * padded] floats in the FP64 slots' space
  if (b->count > 0) {
    bool comm = false;
    for (const PimyParams& q : b->host_pimy) comm |= k.damping * q.betaKcomm != 0.0;
    const dim3 g((b->count + 127) / 128);
    if (comm)
      hipLaunchKernelGGL(pimy_operator32_kernel<true>, g, dim3(128), 0, b->ctx->stream, a, x, b->tangent32, slots);
    else
      hipLaunchKernelGGL(pimy_operator32_kernel<false>, g, dim3(128), 0, b->ctx->stream, a, x, b->tangent32, slots);
  }
  return gather<float>(b, slots, y, done);
}

int pimy_diagonal64(fx_batch* b, const fx_coeffs& k, double* d) {
  if (!b->tangents_stored && !k.initial && k.stiffness != 0.0) {
    b->ctx->error = "operator diagonal: no stored tangents (call fx_store_tangents first)";
    return FX_EINVAL;
  }
  return run_pimy<2, false>(b, batch_args(b, nullptr, k), d, nullptr);
}

int pimy_slots64(fx_batch* b, bool diagonal, const fx_coeffs& k, const double* x, int unit_dir, double* slots) {
  if (!b->tangents_stored && !k.initial && k.stiffness != 0.0) {
    b->ctx->error = "fx_run_operator: no stored tangents (call fx_store_tangents first)";
    return FX_EINVAL;
  }
  PimyBatchArgs a = batch_args(b, x, k);
  a.unit_dir = unit_dir;
  a.slots = slots;
  if (diagonal) launch_operator_slots<true>(b, a);
  else launch_operator_slots<false>(b, a);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

int slots_gather64(fx_batch* b, const double* slots, double* y, int accumulate) {
  return gather<double>(b, slots, y, nullptr, accumulate);
}

int pimy_points(fx_ctx* ctx, int64_t npts, const PimyParams* params, const int32_t* param, const double* C, double* T,
                const double* eps, double* sig, double* D) {
  if (npts <= 0) return FX_OK;
  hipLaunchKernelGGL(pimy_point_kernel, dim3(static_cast<unsigned>((npts + 127) / 128)), dim3(128), 0, ctx->stream,
                     npts, params, param, C, T, eps, sig, D);
  FX_HIP(ctx, hipGetLastError());
  return FX_OK;
}

int pimy_init(fx_ctx* ctx, int64_t npts, const PimyParams* params, const int32_t* param, double* S) {
  if (npts <= 0) return FX_OK;
  hipLaunchKernelGGL(pimy_init_kernel, dim3(static_cast<unsigned>((npts + 255) / 256)), dim3(256), 0, ctx->stream, npts,
                     params, param, S);
  FX_HIP(ctx, hipGetLastError());
  return FX_OK;
}

}  // namespace fx
