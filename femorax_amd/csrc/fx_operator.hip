// hex8 (stdBrick, 2x2x2 Gauss) matrix-free operator kernels for gfx950.
//
// One workgroup per element block (fx_blocks.cpp), one element per thread:
//   gather   the element's 8 nodes, their coordinates and x at the node's equations
//   compute  y_e = sum over the 8 Gauss points of B^T (D_eff B x_e) dvol + cM' (M x_e)
//            (element_operator_apply, src/physics/operators/element_operator.hpp:82-130, with
//            stdBrick's shape functions std_brick.hpp:45-152 and the isotropic D of
//            elastic_isotropic.hpp:40-94), by sum factorisation on the trilinear modes
//            (element_apply_modal)
//   stage    the 24 values in LDS (the reference's slots, assembly.hpp:28-36, without the HBM
//            round trip)
//   reduce   per local node in a fixed (element, corner) order; a node only this block touches
//            is written (stored, or added with accumulate), a node on a block boundary goes to
//            its partial-sum slot, and block_shared_sum_kernel adds a node's slots in block order
// Every block runs in one launch; the scatter is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"
#include "fx_brick_device.hpp"

namespace fx {

// Effective factors of element_operator.hpp:45-56 for an elastic material (D_0 = D_c = D):
// D_eff = (cK + c0 + cc) D; mass factor cM' = cM + cC alphaM.
template <typename T>
struct ElemCoef {
  T lam, mu, rhoM;  // f*lambda, f*mu, cM'*rho
  int consistent;
};


// Diagonal of the element operator, with the cross terms of local DOFs that share an equation
// added to the first of them (element_operator_diagonal, element_operator.hpp:138-209).
template <typename T>
__device__ __forceinline__ void element_diagonal(const T (&x)[8], const T (&y)[8], const T (&z)[8],
                                                 const int (&eq)[24], const ElemCoef<T>& c, T (&out)[24]) {
  int first[24];
  bool merged = false;
#pragma unroll
  for (int a = 0; a < 24; ++a) {
    out[a] = T(0);
    first[a] = a;
  }
  for (int a = 0; a < 24; ++a) {
    if (eq[a] < 0) continue;
    for (int b = 0; b < a; ++b)
      if (eq[b] == eq[a]) {
        first[a] = first[b];
        merged = true;
        break;
      }
  }
  const T l2 = c.lam + T(2) * c.mu;
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    Shape<T> s;
    gauss_shape(x, y, z, gp, s);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const T dx = s.dx[j], dy = s.dy[j], dz = s.dz[j];
      const T m = c.rhoM != T(0) ? c.rhoM * s.dvol * s.N[j] * (c.consistent ? s.N[j] : T(1)) : T(0);
      out[3 * j] += (l2 * dx * dx + c.mu * (dy * dy + dz * dz)) * s.dvol + m;
      out[3 * j + 1] += (l2 * dy * dy + c.mu * (dx * dx + dz * dz)) * s.dvol + m;
      out[3 * j + 2] += (l2 * dz * dz + c.mu * (dx * dx + dy * dy)) * s.dvol + m;
    }
    if (!merged) continue;
    // K_ab for distinct local DOFs a = (J, p), b = (K, q) with one equation:
    // B_J^T D B_K entry (p, q) = lam dJp dKq + mu dJq dKp + (p == q) mu grad_J . grad_K
    for (int a = 0; a < 24; ++a) {
      if (eq[a] < 0) continue;
      for (int b = 0; b < 24; ++b) {
        if (b == a || eq[b] != eq[a]) continue;
        const int J = a / 3, p = a % 3, K = b / 3, q = b % 3;
        const T gJ[3] = {s.dx[J], s.dy[J], s.dz[J]}, gK[3] = {s.dx[K], s.dy[K], s.dz[K]};
        T kab = c.lam * gJ[p] * gK[q] + c.mu * gJ[q] * gK[p];
        if (p == q) kab += c.mu * (gJ[0] * gK[0] + gJ[1] * gK[1] + gJ[2] * gK[2]);
        kab *= s.dvol;
        if (p == q && J != K && c.rhoM != T(0) && c.consistent) kab += c.rhoM * s.dvol * s.N[J] * s.N[K];
        out[first[a]] += kab;
      }
    }
  }
}

// SSPbrick (ssp_brick.hpp) from its constants c[k * P + e] (dNmod, mVol, node weights, hourglass
// vectors, interior's upper triangle): the element operator with its element term
// (element_operator_apply, element_operator.hpp:82-130): fD (mVol Bnot^T D Bnot + Kstab) v plus
// the lumped mass, Kstab v = Mben^T (interior (Mben v)) without the 24 x 24 matrix
// (element_term_reduce / _node, ssp_brick.hpp:964-989). For an elastic material every stiffness
// and damping term has the same tangent, so the point's factor (cK + c0 + cc) and the element
// term's (cK + cC (betaK + betaKinit + betaKcomm), element_term.hpp:34-37) are one fD.
template <typename T>
__device__ __forceinline__ T ssp_at(const double* __restrict__ c, size_t P, size_t e, int k) {
  return static_cast<T>(c[static_cast<size_t>(k) * P + e]);
}

template <typename T>
__device__ __forceinline__ void ssp_element_apply(const double* __restrict__ c, size_t P, size_t e, const T (&v)[24],
                                                  const ElemCoef<T>& cf, T fE, T (&out)[24]) {
  T g[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) g[k] = ssp_at<T>(c, P, e, k);
  const T V = ssp_at<T>(c, P, e, 24);
  T e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const T bx = g[3 * j], by = g[3 * j + 1], bz = g[3 * j + 2];
    const T u0 = v[3 * j], u1 = v[3 * j + 1], u2 = v[3 * j + 2];
    e0 += bx * u0;
    e1 += by * u1;
    e2 += bz * u2;
    e3 += by * u0 + bx * u1;
    e4 += bz * u1 + by * u2;
    e5 += bz * u0 + bx * u2;
  }
  const T lt = cf.lam * (e0 + e1 + e2), m2 = T(2) * cf.mu;
  const T s0 = (m2 * e0 + lt) * V, s1 = (m2 * e1 + lt) * V, s2 = (m2 * e2 + lt) * V;
  const T s3 = cf.mu * e3 * V, s4 = cf.mu * e4 * V, s5 = cf.mu * e5 * V;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const T bx = g[3 * j], by = g[3 * j + 1], bz = g[3 * j + 2];
    out[3 * j] = bx * s0 + by * s3 + bz * s5;
    out[3 * j + 1] = by * s1 + bx * s3 + bz * s4;
    out[3 * j + 2] = bz * s2 + by * s4 + bx * s5;
  }
  if (cf.rhoM != T(0)) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const T m = cf.rhoM * ssp_at<T>(c, P, e, 25 + j);
#pragma unroll
      for (int d = 0; d < 3; ++d) out[3 * j + d] += m * v[3 * j + d];
    }
  }
  if (fE != T(0)) {
    T gam[32], w[12], z[12];
#pragma unroll
    for (int k = 0; k < 32; ++k) gam[k] = ssp_at<T>(c, P, e, kSspGamma + k);
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        T s = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) s += gam[8 * m + i] * v[3 * i + d];
        w[3 * m + d] = s;
        z[3 * m + d] = T(0);
      }
    int idx = kSspInterior;
#pragma unroll
    for (int r = 0; r < 12; ++r)
#pragma unroll
      for (int q = r; q < 12; ++q) {
        const T sv = ssp_at<T>(c, P, e, idx++);
        z[r] += sv * w[q];
        if (q != r) z[q] += sv * w[r];
      }
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        T s = 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) s += gam[8 * m + j] * z[3 * m + d];
        out[3 * j + d] += fE * s;
      }
  }
}

// interior(r, q) from the packed upper triangle
template <typename T>
__device__ __forceinline__ T ssp_interior(const double* __restrict__ c, size_t P, size_t e, int r, int q) {
  if (r > q) {
    const int t = r;
    r = q;
    q = t;
  }
  return ssp_at<T>(c, P, e, kSspInterior + 12 * r - r * (r - 1) / 2 + (q - r));
}

// Kstab(a, b) = sum_gh gamma_g(i) interior(3g + d, 3h + e) gamma_h(k) (element_term_entry,
// ssp_brick.hpp:991-1005)
template <typename T>
__device__ __forceinline__ T ssp_term_entry(const double* __restrict__ c, size_t P, size_t e, int a, int b) {
  const int i = a / 3, d = a % 3, k = b / 3, f = b % 3;
  T s = 0;
  for (int g = 0; g < 4; ++g) {
    const T gi = ssp_at<T>(c, P, e, kSspGamma + 8 * g + i);
    for (int h = 0; h < 4; ++h)
      s += gi * ssp_interior<T>(c, P, e, 3 * g + d, 3 * h + f) * ssp_at<T>(c, P, e, kSspGamma + 8 * h + k);
  }
  return s;
}

// Diagonal of the SSPbrick operator, the cross terms of local DOFs that share an equation added to
// the first of them (element_operator_diagonal, element_operator.hpp:138-209; the lumped mass has
// no off-diagonal term).
template <typename T>
__device__ __forceinline__ void ssp_element_diagonal(const double* __restrict__ c, size_t P, size_t e,
                                                     const int (&eq)[24], const ElemCoef<T>& cf, T fE, T (&out)[24]) {
  int first[24];
  bool merged = false;
#pragma unroll
  for (int a = 0; a < 24; ++a) first[a] = a;
  for (int a = 0; a < 24; ++a) {
    if (eq[a] < 0) continue;
    for (int b = 0; b < a; ++b)
      if (eq[b] == eq[a]) {
        first[a] = first[b];
        merged = true;
        break;
      }
  }
  const T V = ssp_at<T>(c, P, e, 24), l2 = cf.lam + T(2) * cf.mu;
#pragma unroll 1
  for (int j = 0; j < 8; ++j) {
    const T dx = ssp_at<T>(c, P, e, 3 * j), dy = ssp_at<T>(c, P, e, 3 * j + 1), dz = ssp_at<T>(c, P, e, 3 * j + 2);
    const T m = cf.rhoM != T(0) ? cf.rhoM * ssp_at<T>(c, P, e, 25 + j) : T(0);
    out[3 * j] = (l2 * dx * dx + cf.mu * (dy * dy + dz * dz)) * V + m;
    out[3 * j + 1] = (l2 * dy * dy + cf.mu * (dx * dx + dz * dz)) * V + m;
    out[3 * j + 2] = (l2 * dz * dz + cf.mu * (dx * dx + dy * dy)) * V + m;
  }
  if (fE != T(0))
    for (int a = 0; a < 24; ++a) out[a] += fE * ssp_term_entry<T>(c, P, e, a, a);
  if (!merged) return;
  for (int a = 0; a < 24; ++a) {
    if (eq[a] < 0) continue;
    for (int b = 0; b < 24; ++b) {
      if (b == a || eq[b] != eq[a]) continue;
      const int J = a / 3, p = a % 3, K = b / 3, q = b % 3;
      const T gJ[3] = {ssp_at<T>(c, P, e, 3 * J), ssp_at<T>(c, P, e, 3 * J + 1), ssp_at<T>(c, P, e, 3 * J + 2)};
      const T gK[3] = {ssp_at<T>(c, P, e, 3 * K), ssp_at<T>(c, P, e, 3 * K + 1), ssp_at<T>(c, P, e, 3 * K + 2)};
      T kab = cf.lam * gJ[p] * gK[q] + cf.mu * gJ[q] * gK[p];
      if (p == q) kab += cf.mu * (gJ[0] * gK[0] + gJ[1] * gK[1] + gJ[2] * gK[2]);
      kab *= V;
      if (fE != T(0)) kab += fE * ssp_term_entry<T>(c, P, e, a < b ? a : b, a < b ? b : a);
      out[first[a]] += kab;
    }
  }
}

template <typename T>
__device__ __forceinline__ T load_coord(const double* c64, const float* c32, size_t i) {
  if constexpr (sizeof(T) == 8) return c64[i];
  else return c32[i];
}

// The block kernel: y (or the node-space ynode) += K_eff x over the block's elements. A block
// whose local nodes fit (kStageNodes) first loads them, in block order, into the LDS its slots
// use afterwards: each node's coordinates, equations and x once (coalesced, where an element-wise
// gather would fetch them for each of up to eight elements); an element then reads its corners
// there. Larger blocks gather per element (GATHER: a second launch over the same grid, made
// only when the batch has such blocks; each launch skips the other's blocks). The values are the
// same either way, so are the bits. Registers are held to two waves per SIMD for FP64 (the
// staged FP64 element work fits in 228), four for FP32.
template <typename T, bool DIAG, bool SSP, bool GATHER = false>
__global__ __launch_bounds__(kBlockElems, sizeof(T) == 8 ? 2 : 4) void hex8_block_kernel(KernelArgs<T> a,
                                                                                         T* __restrict__ partial) {
  extern __shared__ unsigned char smem_raw[];
  T* slots = reinterpret_cast<T*>(smem_raw);  // [24][kBlockElems]
  if (a.done && *a.done) return;
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int ln0 = a.ln_start[b], ln1 = a.ln_start[b + 1];
  if ((ln1 - ln0 > kStageNodes) != GATHER) return;  // the other launch's block (uniform)
  const int e = a.elems[static_cast<size_t>(b) * kBlockElems + t];
  const size_t tot = static_cast<size_t>(a.ln_total);
  T X[8], Y[8], Z[8], V[24], out[24];
  int eqe[24];
  // the element's operator values (or diagonal) from its corners; vl(i) reloads V[i]
  auto element = [&](auto vl) {
    const DevParam& p = a.params[a.param[e]];
    double fD, cM;
    operator_factors(p, a.k, fD, cM);
    ElemCoef<T> c;
    c.lam = static_cast<T>(fD * p.lam);
    c.mu = static_cast<T>(fD * p.mu);
    c.rhoM = static_cast<T>(cM * p.rho);
    c.consistent = p.mass_type == 0;
    if constexpr (SSP) {
      // the constants in block order: coalesced loads
      const size_t at = static_cast<size_t>(b) * kBlockElems + t;
      if constexpr (DIAG) ssp_element_diagonal<T>(a.ssp, a.ssp_stride, at, eqe, c, static_cast<T>(fD), out);
      else ssp_element_apply<T>(a.ssp, a.ssp_stride, at, V, c, static_cast<T>(fD), out);
    } else {
      if constexpr (DIAG) element_diagonal(X, Y, Z, eqe, c, out);
      else element_apply_modal<T>(X, Y, Z, V, c.lam, c.mu, c.rhoM, c.consistent != 0, out, vl);
    }
  };
  if constexpr (!GATHER) {
    // staging: x, y, z, then the three x values (or, for the diagonal, equations) per node
    T* sxyz = slots;
    T* sv = slots + 3 * kStageNodes;
    int* seq = reinterpret_cast<int*>(sv);
    for (int i = t; i < ln1 - ln0; i += kBlockElems) {
      const size_t ln = static_cast<size_t>(ln0) + i;
      if constexpr (!SSP)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if constexpr (sizeof(T) == 8) sxyz[d * kStageNodes + i] = a.ln_xyz64[d * tot + ln];
          else sxyz[d * kStageNodes + i] = a.ln_xyz32[d * tot + ln];
        }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int q = a.ln_eq[d * tot + ln];
        if constexpr (DIAG) seq[d * kStageNodes + i] = q;
        else sv[d * kStageNodes + i] = q >= 0 ? a.x[q] : T(0);
      }
    }
    __syncthreads();
    if (e >= 0) {
      const uint16_t* lc = a.lcorner + static_cast<size_t>(b) * 8 * kBlockElems + t;
      int li[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        li[k] = lc[k * kBlockElems];
        if constexpr (!SSP) {
          X[k] = sxyz[li[k]];
          Y[k] = sxyz[kStageNodes + li[k]];
          Z[k] = sxyz[2 * kStageNodes + li[k]];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if constexpr (DIAG) eqe[3 * k + d] = seq[d * kStageNodes + li[k]];
          else V[3 * k + d] = sv[d * kStageNodes + li[k]];
        }
      }
      element([&](int i) { return sv[(i % 3) * kStageNodes + li[i / 3]]; });
    }
    __syncthreads();  // the staging area becomes the slots
  } else if (e >= 0) {  // per-element gathers
    const size_t P = static_cast<size_t>(a.padded);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int n = a.node[k * P + e];
      if constexpr (!SSP) {
        X[k] = load_coord<T>(a.coords64, a.coords32, n);
        Y[k] = load_coord<T>(a.coords64, a.coords32, static_cast<size_t>(a.nnode) + n);
        Z[k] = load_coord<T>(a.coords64, a.coords32, 2 * static_cast<size_t>(a.nnode) + n);
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int q = a.node_eq[3 * n + d];
        eqe[3 * k + d] = q;
        if constexpr (!DIAG) V[3 * k + d] = q >= 0 ? a.x[q] : T(0);
      }
    }
    element([&](int i) { return V[i]; });
  }
  if (e >= 0) {
#pragma unroll
    for (int k = 0; k < 24; ++k) slots[k * kBlockElems + t] = out[k];
  }
  __syncthreads();
  const uint16_t* list = a.slot_list + static_cast<size_t>(b) * 8 * kBlockElems;
  for (int ln = ln0 + t; ln < ln1; ln += kBlockElems) {
    const uint32_t sc = a.ln_slots[ln];
    const int s0 = sc & 0xffff, ns = sc >> 16;
    T acc0 = 0, acc1 = 0, acc2 = 0;
    // the node's (element, corner) slots, eight at a time: their list entries are loaded before
    // the LDS reads use them (one round of memory latency per eight, not one per slot)
    for (int j0 = 0; j0 < ns; j0 += 8) {
      int sl[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) sl[j] = j0 + j < ns ? list[s0 + j0 + j] : -1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (sl[j] < 0) break;
        const int el = sl[j] >> 3, corner = sl[j] & 7;
        acc0 += slots[(3 * corner) * kBlockElems + el];
        acc1 += slots[(3 * corner + 1) * kBlockElems + el];
        acc2 += slots[(3 * corner + 2) * kBlockElems + el];
      }
    }
    const T acc[3] = {acc0, acc1, acc2};
    const int part = a.ln_part[ln];
    if (part >= 0) {  // a node other blocks touch too: this block's partial sum
#pragma unroll
      for (int d = 0; d < 3; ++d) partial[3 * static_cast<size_t>(part) + d] = acc[d];
      continue;
    }
    if (a.ynode) {  // tied mesh: node space, contracted afterwards
      const int n = static_cast<int>(a.lnode[ln]);
#pragma unroll
      for (int d = 0; d < 3; ++d) a.ynode[3 * static_cast<size_t>(n) + d] = acc[d];
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int q = a.ln_eq[d * tot + ln];
        if (q < 0) continue;
        a.y[q] = a.accumulate ? a.y[q] + acc[d] : acc[d];
      }
    }
  }
}

// The nodes several blocks touch: their blocks' partial sums added in block order.
template <typename T>
__global__ void block_shared_sum_kernel(int32_t nshared, const int32_t* __restrict__ sh_node,
                                        const int32_t* __restrict__ sh_start, const T* __restrict__ partial,
                                        const int32_t* __restrict__ node_eq, T* __restrict__ y, T* __restrict__ ynode,
                                        int accumulate, const int32_t* done) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshared) return;
  const int n = sh_node[i];
  T acc[3] = {T(0), T(0), T(0)};
  for (int s = sh_start[i]; s < sh_start[i + 1]; ++s)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] += partial[3 * static_cast<size_t>(s) + d];
  if (ynode) {
#pragma unroll
    for (int d = 0; d < 3; ++d) ynode[3 * static_cast<size_t>(n) + d] = acc[d];
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int q = node_eq[3 * n + d];
      if (q < 0) continue;
      y[q] = accumulate ? y[q] + acc[d] : acc[d];
    }
  }
}

// One element per thread: element e's 24 operator values (or its diagonal) into slots[a * padded
// + e], for every element e < count (launch_operator with emit's Out::Slots,
// operator_kernels.hpp:75-141, assembly.hpp:28-36). unit_dir >= 0 takes the rigid unit
// translation in that direction as the input at every DOF, fixed ones included (:96-98).
template <bool DIAG, bool SSP>
__global__ __launch_bounds__(128) void element_slots_kernel(KernelArgs<double> a, int32_t count, int unit_dir,
                                                             double* __restrict__ slots) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  double X[8], Y[8], Z[8], V[24], out[24];
  int eqe[24];
  const size_t P = static_cast<size_t>(a.padded);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int n = a.node[k * P + e];
    X[k] = a.coords64[n];
    Y[k] = a.coords64[static_cast<size_t>(a.nnode) + n];
    Z[k] = a.coords64[2 * static_cast<size_t>(a.nnode) + n];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int q = a.node_eq[3 * n + d];
      eqe[3 * k + d] = q;
      if (unit_dir >= 0) V[3 * k + d] = d == unit_dir ? 1.0 : 0.0;
      else V[3 * k + d] = (!DIAG && q >= 0) ? a.x[q] : 0.0;
    }
  }
  const DevParam& p = a.params[a.param[e]];
  double fD, cM;
  operator_factors(p, a.k, fD, cM);
  ElemCoef<double> c;
  c.lam = fD * p.lam;
  c.mu = fD * p.mu;
  c.rhoM = cM * p.rho;
  c.consistent = p.mass_type == 0;
  if constexpr (SSP) {
    const size_t at = a.ssp_pos ? static_cast<size_t>(a.ssp_pos[e]) : static_cast<size_t>(e);
    if constexpr (DIAG) ssp_element_diagonal<double>(a.ssp, a.ssp_stride, at, eqe, c, fD, out);
    else ssp_element_apply<double>(a.ssp, a.ssp_stride, at, V, c, fD, out);
  } else {
    if constexpr (DIAG) element_diagonal(X, Y, Z, eqe, c, out);
    else element_apply_modal<double>(X, Y, Z, V, c.lam, c.mu, c.rhoM, c.consistent != 0, out);
  }
#pragma unroll
  for (int k = 0; k < 24; ++k) slots[k * P + e] = out[k];
}

// y[q] (+)= sum of the node-space values of q's node DOFs, in ascending DOF order.
template <typename T>
__global__ void contract_kernel(int32_t neq, const int32_t* __restrict__ start, const int32_t* __restrict__ list,
                                const T* __restrict__ ynode, T* __restrict__ y, int accumulate, const int32_t* done) {
  if (done && *done) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= neq) return;
  T s = 0;
  for (int j = start[q]; j < start[q + 1]; ++j) s += ynode[list[j]];
  y[q] = accumulate ? y[q] + s : s;
}

template <typename T, bool DIAG>
int launch_operator(fx_batch* b, const KernelArgs<T>& args) {
  hipStream_t st = b->ctx->stream;
  const size_t smem = sizeof(T) * 24 * kBlockElems;
  T* partial;
  if constexpr (sizeof(T) == 8) partial = b->partial64;
  else partial = b->partial32;
  if (b->blk.nshared > 0 && !partial) {
    b->ctx->error = "operator: no partial-sum buffer for this precision";
    return FX_EINVAL;
  }
  if (b->blk.nblocks > 0) {
    const dim3 grid(b->blk.nblocks);
    if (b->ssp)
      hipLaunchKernelGGL((hex8_block_kernel<T, DIAG, true>), grid, dim3(kBlockElems), smem, st, args, partial);
    else
      hipLaunchKernelGGL((hex8_block_kernel<T, DIAG, false>), grid, dim3(kBlockElems), smem, st, args, partial);
    if (b->blk.max_ln > kStageNodes) {  // blocks whose local nodes do not fit the staging area
      if (b->ssp)
        hipLaunchKernelGGL((hex8_block_kernel<T, DIAG, true, true>), grid, dim3(kBlockElems), smem, st, args,
                           partial);
      else
        hipLaunchKernelGGL((hex8_block_kernel<T, DIAG, false, true>), grid, dim3(kBlockElems), smem, st, args,
                           partial);
    }
  }
  if (b->blk.nshared > 0)
    hipLaunchKernelGGL(block_shared_sum_kernel<T>, dim3((b->blk.nshared + 255) / 256), dim3(256), 0, st,
                       b->blk.nshared, b->blk.sh_node, b->blk.sh_start, partial, args.node_eq, args.y, args.ynode,
                       args.accumulate, args.done);
  if (b->tied) {
    const int nb = (b->neq + 255) / 256;
    hipLaunchKernelGGL(contract_kernel<T>, dim3(nb), dim3(256), 0, st, b->neq, b->con_start, b->con_list,
                       args.ynode, args.y, args.accumulate, args.done);
  }
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

template <typename T>
KernelArgs<T> operator_args(fx_batch* b, const fx_coeffs& k, const T* x, T* y, int accumulate,
                            const int32_t* done = nullptr) {
  KernelArgs<T> a{};
  a.done = done;
  a.k = k;
  a.padded = b->padded;
  a.nnode = b->nnode;
  a.node = b->node;
  a.param = b->param;
  a.node_eq = b->node_eq;
  a.coords64 = b->coords;
  a.coords32 = b->coords32;
  a.params = b->params;
  a.elems = b->blk.elems;
  a.ln_start = b->blk.ln_start;
  a.lnode = b->blk.lnode;
  a.ln_slots = b->blk.ln_slots;
  a.slot_list = b->blk.slot_list;
  a.ln_part = b->blk.ln_part;
  a.lcorner = b->blk.lcorner;
  a.ln_total = b->blk.ln_total;
  a.ln_eq = b->blk.ln_eq;
  a.ln_xyz64 = b->blk.ln_xyz;
  a.ln_xyz32 = b->blk.ln_xyz32;
  a.x = x;
  a.y = y;
  a.accumulate = accumulate;
  a.ssp = b->ssp_elem;
  a.ssp_stride = b->lattice ? b->padded : static_cast<int64_t>(b->blk.nblocks) * kBlockElems;
  a.ssp_pos = b->lattice ? nullptr : b->blk.pos;
  if (b->tied) {
    if constexpr (sizeof(T) == 8) a.ynode = b->ynode;
    else a.ynode = b->ynode32;
  }
  return a;
}

template <typename T>
__global__ void zero_orphans_kernel(int32_t n, const int32_t* __restrict__ list, T* __restrict__ y, const int32_t* done) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[list[i]] = T(0);
}

template <typename T>
int zero_orphans(fx_batch* b, T* y, const int32_t* done) {
  if (b->norphans == 0) return FX_OK;
  hipLaunchKernelGGL(zero_orphans_kernel<T>, dim3((b->norphans + 255) / 256), dim3(256), 0, b->ctx->stream, b->norphans,
                     b->orphans, y, done);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

// y += c m x (or c m on the diagonal): add_nodal_mass (internal_force.cpp:442-448).
template <typename T>
__global__ void nodal_mass_kernel(int32_t n, T c, const T* __restrict__ m, const T* __restrict__ x, T* __restrict__ y,
                                  int diagonal, const int32_t* done) {
  if (done && *done) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) y[q] += c * m[q] * (diagonal ? T(1) : x[q]);
}

// The nodal masses' terms of the operator (cM m x, then cC am x) or of its diagonal
// (internal_force.cpp:571-572, 597-598).
template <typename T>
int add_nodal(fx_batch* b, const fx_coeffs& k, const T* x, T* y, int diagonal, const int32_t* done) {
  const T* m;
  const T* am;
  if constexpr (sizeof(T) == 8) {
    m = b->nodal_m;
    am = b->nodal_am;
  } else {
    m = b->nodal_m32;
    am = b->nodal_am32;
  }
  const dim3 g((b->neq + 255) / 256);
  if (m && k.mass != 0.0)
    hipLaunchKernelGGL(nodal_mass_kernel<T>, g, dim3(256), 0, b->ctx->stream, b->neq, static_cast<T>(k.mass), m, x, y,
                       diagonal, done);
  if (am && k.damping != 0.0)
    hipLaunchKernelGGL(nodal_mass_kernel<T>, g, dim3(256), 0, b->ctx->stream, b->neq, static_cast<T>(k.damping), am, x,
                       y, diagonal, done);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

int nodal_damping_force64(fx_batch* b, const double* v, double* f) {
  return add_nodal<double>(b, fx_coeffs{0.0, 1.0, 0.0, 0}, v, f, 0, nullptr);
}

int elements_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y, const int32_t* done) {
  if (b->pimy) return pimy_apply64(b, k, x, y, done);
  const int rc = b->lattice ? launch_box<double, false>(b, k, x, y, 0, done)
                            : launch_operator<double, false>(b, operator_args<double>(b, k, x, y, 0, done));
  return rc ? rc : zero_orphans(b, y, done);
}
int operator_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y, const int32_t* done) {
  if (int rc = elements_apply64(b, k, x, y, done)) return rc;
  return (b->nodal_m || b->nodal_am) ? add_nodal<double>(b, k, x, y, 0, done) : FX_OK;
}
int operator_apply32(fx_batch* b, const fx_coeffs& k, const float* x, float* y, const int32_t* done) {
  int rc;
  if (b->pimy) {
    rc = pimy_apply32(b, k, x, y, done);
  } else {
    rc = b->lattice ? launch_box<float, false>(b, k, x, y, 0, done)
                    : launch_operator<float, false>(b, operator_args<float>(b, k, x, y, 0, done));
    if (!rc) rc = zero_orphans(b, y, done);
  }
  if (rc) return rc;
  return (b->nodal_m32 || b->nodal_am32) ? add_nodal<float>(b, k, x, y, 0, done) : FX_OK;
}

int element_slots64(fx_batch* b, bool diagonal, const fx_coeffs& k, const double* x, int unit_dir, double* slots) {
  if (!b->node || !b->node_eq || !b->coords || !b->param) {
    b->ctx->error = "fx_run_operator: the batch keeps no element connectivity on the device (create it with "
                    "FX_BATCH_SLOTS)";
    return FX_EINVAL;
  }
  if (b->count == 0) return FX_OK;
  const KernelArgs<double> a = operator_args<double>(b, k, x, static_cast<double*>(nullptr), 0);
  const dim3 g((b->count + 127) / 128);
  if (b->ssp && !b->ssp_elem) {
    b->ctx->error = "fx_run_operator: the SSPbrick batch keeps no element constants (create it with FX_BATCH_SLOTS)";
    return FX_EINVAL;
  }
  auto* kern = diagonal ? (b->ssp ? element_slots_kernel<true, true> : element_slots_kernel<true, false>)
                        : (b->ssp ? element_slots_kernel<false, true> : element_slots_kernel<false, false>);
  hipLaunchKernelGGL(kern, g, dim3(128), 0, b->ctx->stream, a, b->count, unit_dir, slots);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}
__global__ void add_at_kernel(int32_t n, const int32_t* __restrict__ idx, const double* __restrict__ val,
                              double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[idx[i]] += val[i];
}

int operator_diagonal64(fx_batch* b, const fx_coeffs& k, double* d) {
  if (int rc = elements_diagonal64(b, k, d)) return rc;
  return (b->nodal_m || b->nodal_am) ? add_nodal<double>(b, k, static_cast<const double*>(nullptr), d, 1, nullptr)
                                     : FX_OK;
}

int elements_diagonal64(fx_batch* b, const fx_coeffs& k, double* d) {
  if (b->pimy) return pimy_diagonal64(b, k, d);
  const int rc = b->lattice
                     ? launch_box<double, true>(b, k, static_cast<const double*>(nullptr), d, 0, nullptr)
                     : launch_operator<double, true>(b, operator_args<double>(b, k, static_cast<const double*>(nullptr), d, 0));
  if (rc) return rc;
  if (!b->tie_eqs.empty()) {
    // a tied lattice: the cross terms K_ab + M_ab of the DOF pairs an element ties together, for
    // this call's coefficients, added to their equations
    std::fill(b->tie_host.begin(), b->tie_host.end(), 0.0);
    size_t u = 0;
    for (size_t i = 0; i < b->tie_q.size(); ++i) {
      while (b->tie_eqs[u] != b->tie_q[i]) ++u;
      const DevParam& p = b->host_params[b->tie_p[i]];
      double fD, cM;
      operator_factors(p, k, fD, cM);
      b->tie_host[u] += fD * (p.lam * b->tie_kl[i] + p.mu * b->tie_km[i]) + (p.mass_type == 0 ? cM * p.rho * b->tie_mc[i] : 0.0);
    }
    const int32_t n = static_cast<int32_t>(b->tie_eqs.size());
    FX_HIP(b->ctx, hipMemcpyAsync(b->tie_val, b->tie_host.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                  b->ctx->stream));
    hipLaunchKernelGGL(add_at_kernel, dim3((n + 255) / 256), dim3(256), 0, b->ctx->stream, n, b->tie_idx, b->tie_val, d);
    FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));  // the host vector is reused by the next call
  }
  return zero_orphans(b, d, static_cast<const int32_t*>(nullptr));
}

}  // namespace fx

namespace fx {

template <typename T>
__global__ void gather_perm_kernel(int32_t n, const int32_t* __restrict__ perm, const T* __restrict__ src,
                                   T* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
template <typename T>
__global__ void scatter_perm_kernel(int32_t n, const int32_t* __restrict__ perm, const T* __restrict__ src,
                                    T* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}

template <typename T>
int to_internal(fx_batch* b, const T* ext, T* in) {
  if (b->neq == 0) return FX_OK;
  if (!b->internal) {
    if (ext != in) FX_HIP(b->ctx, hipMemcpyAsync(in, ext, sizeof(T) * b->neq, hipMemcpyDeviceToDevice, b->ctx->stream));
    return FX_OK;
  }
  hipLaunchKernelGGL(gather_perm_kernel<T>, dim3((b->neq + 255) / 256), dim3(256), 0, b->ctx->stream, b->neq, b->perm,
                     ext, in);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}
template <typename T>
int to_external(fx_batch* b, const T* in, T* ext) {
  if (b->neq == 0) return FX_OK;
  if (!b->internal) {
    if (ext != in) FX_HIP(b->ctx, hipMemcpyAsync(ext, in, sizeof(T) * b->neq, hipMemcpyDeviceToDevice, b->ctx->stream));
    return FX_OK;
  }
  hipLaunchKernelGGL(scatter_perm_kernel<T>, dim3((b->neq + 255) / 256), dim3(256), 0, b->ctx->stream, b->neq, b->perm,
                     in, ext);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}
template int to_internal<double>(fx_batch*, const double*, double*);
template int to_internal<float>(fx_batch*, const float*, float*);
template int to_external<double>(fx_batch*, const double*, double*);
template int to_external<float>(fx_batch*, const float*, float*);

namespace {
// The drop-in apply in place on reference-numbered vectors: lattice batches with reference words,
// no nodal masses (their arrays are in the solver numbering)
bool direct_ref(const fx_batch* b) {
  return b->lattice && b->internal && b->words_ref && !b->pimy && !b->nodal_m && !b->nodal_am;
}

template <typename T>
__global__ void zero_list_kernel(int32_t n, const int32_t* __restrict__ list, T* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[list[i]] = T(0);
}

template <typename T>
int apply_ref(fx_batch* b, const fx_coeffs& k, const T* x, T* y) {
  if (int rc = launch_box<T, false>(b, k, x, y, 0, nullptr, 1)) return rc;
  if (b->norphans > 0)
    hipLaunchKernelGGL(zero_list_kernel<T>, dim3((b->norphans + 255) / 256), dim3(256), 0, b->ctx->stream, b->norphans,
                       b->orphans_ref, y);
  FX_HIP(b->ctx, hipGetLastError());
  if (!b->ctx->stream_ordered) FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}

template <typename T, typename Op>
int external_call(fx_batch* b, const T* x, T* y, T* xi, T* yi, Op op) {
  if (b->internal) {
    if (x) {
      if (int rc = to_internal<T>(b, x, xi)) return rc;
    }
    if (int rc = op(x ? xi : nullptr, yi)) return rc;
    if (int rc = to_external<T>(b, yi, y)) return rc;
  } else {
    if (int rc = op(x, y)) return rc;
  }
  if (!b->ctx->stream_ordered) FX_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return FX_OK;
}
}  // namespace

__global__ void add_perm_kernel(int32_t n, const int32_t* __restrict__ perm, const double* __restrict__ src,
                                double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm ? perm[i] : i] += src[i];
}

// y (reference numbering) += this batch's operator or diagonal without the nodal masses, computed
// on the fast path into the batch's scratch (fx_run_operator's FX_OUT_PLAIN / FX_OUT_ATOMIC).
int add_elements64(fx_batch* b, bool diagonal, const fx_coeffs& k, const double* x, double* y) {
  if (b->neq == 0) return FX_OK;
  if (!diagonal && b->lattice && b->internal && b->words_ref)  // in place: y += the sums (orphans get none)
    return launch_box<double, false>(b, k, x, y, 1, nullptr, 1);
  double* xi = b->work + 6 * static_cast<size_t>(b->neq);
  double* yi = b->work + 7 * static_cast<size_t>(b->neq);
  if (diagonal) {
    if (int rc = elements_diagonal64(b, k, yi)) return rc;
  } else {
    const double* xs = x;
    if (b->internal) {
      if (int rc = to_internal<double>(b, x, xi)) return rc;
      xs = xi;
    }
    if (int rc = elements_apply64(b, k, xs, yi)) return rc;
  }
  hipLaunchKernelGGL(add_perm_kernel, dim3((b->neq + 255) / 256), dim3(256), 0, b->ctx->stream, b->neq,
                     b->internal ? b->perm : static_cast<const int32_t*>(nullptr), yi, y);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

int external_apply64(fx_batch* b, const fx_coeffs& k, const double* x, double* y) {
  if (direct_ref(b)) return apply_ref<double>(b, k, x, y);
  double* xi = b->work + 6 * static_cast<size_t>(b->neq);
  double* yi = b->work + 7 * static_cast<size_t>(b->neq);
  return external_call<double>(b, x, y, xi, yi,
                               [&](const double* xx, double* yy) { return operator_apply64(b, k, xx, yy); });
}
int external_apply32(fx_batch* b, const fx_coeffs& k, const float* x, float* y) {
  if (direct_ref(b)) return apply_ref<float>(b, k, x, y);
  float* xi = b->work32 + 7 * static_cast<size_t>(b->neq);
  float* yi = b->work32 + 8 * static_cast<size_t>(b->neq);
  return external_call<float>(b, x, y, xi, yi, [&](const float* xx, float* yy) { return operator_apply32(b, k, xx, yy); });
}
int external_diagonal64(fx_batch* b, const fx_coeffs& k, double* d) {
  double* yi = b->work + 7 * static_cast<size_t>(b->neq);
  return external_call<double>(b, static_cast<const double*>(nullptr), d, static_cast<double*>(nullptr), yi,
                               [&](const double*, double* yy) { return operator_diagonal64(b, k, yy); });
}

}  // namespace fx
