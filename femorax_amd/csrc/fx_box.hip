// Structured-lattice hex8 operator: one workgroup per tile of 8 x 8 x 4 axis-aligned bricks.
//
// The element operator of stdBrick (element_operator_apply, element_operator.hpp:82-130;
// shp3d at the 2x2x2 Gauss points, std_brick.hpp:45-152) for an axis-aligned box is a tensor
// product: the shape function derivative along x at Gauss point (q1, q2, q3) is
// +-1/hx * I[q2][j] * I[q3][k] with I the 2x2 matrix of the 1D linear basis at the two Gauss
// points. Sum factorisation then evaluates B x at the 8 points with 4 differences and two
// 2-point interpolations per (component, direction) and B^T sigma with the transposed steps,
// about 600 FP64 operations per element instead of the ~6000 of the generic form. D_eff and
// dvol are per parameter set and per lattice; the geometry is three numbers (hx, hy, hz).
//
// Memory per tile: one 32-bit word per lattice node (first equation, free-DOF mask, shared
// flag), a 16-bit parameter index per cell, x at the tile's nodes, y written once per node.
// The element results are summed per node inside the tile with lane shuffles (x and y within a
// wave, one wave per z-plane of cells) and one LDS exchange between neighbouring planes, in a
// fixed order. One launch runs every tile with no inter-workgroup dependency: a node that only
// its tile touches gets its final y there; for a node on a tile face the tile stores its sum in
// its own partial-sum slot, and a second kernel adds a shared node's slots in tile parity-colour
// order (tiles of one colour share no node). The sums are deterministic: the same bits for
// every run, tile placement and co-runner.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "fx_internal.hpp"
#include "fx_kernels.hpp"

namespace fx {

template <typename T>
struct BoxConst {
  T w0[3];   // 8 kappa_d, kappa_d = 1 / (4 h_d): the physical gradient of a linear mode is kappa_d (Hv)
  T w1[3];   // 8/3 kappa_d
  T kap[3];  // kappa_d
  T kk[3];   // 8/9 kappa_d^2
  T kks;     // 8/9 sum_d kappa_d^2
};

template <typename T>
struct BoxArgs {
  const TileHeader* tiles;
  const uint32_t* words;
  const int32_t* slots;  // per tile node: partial-sum slot of a shared node
  const uint16_t* cell_param;
  const DevParam* params;
  fx_coeffs k;
  double hx, hy, hz;
  BoxConst<T> K;
  const T* x;
  T* y;
  T* partial;  // [slots][3]: the tiles' sums at their shared nodes
  int ntiles;
  int accumulate;
  int use_interior;  // regular tiles' interior nodes addressed from their base (internal numbering only)
  const int32_t* done;
  const double* ssp;  // SSPbrick: per parameter set, the modal stabilisation and diag(K_e) (kSspBox each)
  // rectilinear lattices: (kappa, h) of each cell interval along x, y, z
  const double* kx;
  const double* ky;
  const double* kz;
};

// The element operator in the modal basis of the trilinear field.
//
// With corner values v[q], q = i + 2j + 4k, the field is u = sum_m a_m phi_m over the modes
// phi_m = xi^mx eta^my zeta^mz (m = mx + 2my + 4mz), a = H v / 8 with H the 8 x 8 Hadamard
// matrix (one sum/difference butterfly per axis). On an axis-aligned box the gradient of each
// mode is a polynomial whose terms are orthogonal under the 2 x 2 x 2 Gauss rule, which
// integrates every term of the stdBrick stiffness and consistent mass exactly. The energy then
// splits into seven small independent pieces: the constant-strain part (the three linear modes
// of each component), one part per axis p with the bilinear modes that vary along p, and the
// trilinear modes (diagonal), and the consistent mass is diagonal in this basis. K_e v is the
// transposed butterflies of the modal gradient g = dE/d(Hv): about 250 FP64 operations per
// element where B^T D B at the Gauss points takes about 600, with the same result up to rounding.

template <typename T>
__device__ __forceinline__ void hadamard8(T (&v)[8]) {
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
__device__ __forceinline__ void hadamard8_transposed(T (&g)[8]) {
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

// y[c][q] = (K_e v)[c][q] (+ the mass term), v[c][q] from the tile's node array in LDS (xs[c][ln],
// ln[q] the tile node of corner q). lam, mu carry dvol and the operator factor; mc = cM rho dvol
// (MASS only; lumped: the reference's lumped mass, else consistent).
// The constant-strain part (the three linear modes of each component): F[d][c] = kappa_d b[c][e_d],
// g[c][e_d] = 8 kappa_d sigma_dc.
template <typename T>
__device__ __forceinline__ void box_constant_strain(const T (&b)[3][8], T lam, T mu, const BoxConst<T>& K,
                                                    T (&y)[3][8]) {
  const T m2 = mu + mu;
  {
    T F[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 3; ++c) F[d][c] = K.kap[d] * b[c][1 << d];
    const T ltr = lam * (F[0][0] + F[1][1] + F[2][2]);
#pragma unroll
    for (int d = 0; d < 3; ++d) y[d][1 << d] = K.w0[d] * (ltr + m2 * F[d][d]);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = d + 1; c < 3; ++c) {
        const T s = mu * (F[d][c] + F[c][d]);
        y[c][1 << d] = K.w0[d] * s;
        y[d][1 << c] = K.w0[c] * s;
      }
  }
}

template <typename T, bool MASS, int NODES>
__device__ __forceinline__ void box_apply(const T* xs, const int (&ln)[8], T lam, T mu, T mc, bool lumped,
                                          const BoxConst<T>& K, T (&y)[3][8]) {
  T b[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int q = 0; q < 8; ++q) b[c][q] = xs[c * NODES + ln[q]];
    hadamard8(b[c]);
  }
  const T m2 = mu + mu;
  box_constant_strain(b, lam, mu, K, y);
  // the bilinear modes that vary along p: rows d1, d2 (the other axes) of the gradient's p-linear
  // term, F[d][c] = kappa_d b[c][e_d | e_p]; g[c][e_d | e_p] gets 8/3 kappa_d sigma_dc. Mode
  // (d, p) hears from p and from d; the smaller of the two assigns, the other adds.
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const int d1 = p == 0 ? 1 : 0, d2 = p == 2 ? 1 : 2;
    const int n1 = (1 << d1) | (1 << p), n2 = (1 << d2) | (1 << p);
    T F1[3], F2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      F1[c] = K.kap[d1] * b[c][n1];
      F2[c] = K.kap[d2] * b[c][n2];
    }
    const T ltr = lam * (F1[d1] + F2[d2]);
    const T s12 = mu * (F1[d2] + F2[d1]);
    const T r1[3] = {d1 == 0 ? ltr + m2 * F1[0] : (p == 0 ? mu * F1[0] : s12),
                     d1 == 1 ? ltr + m2 * F1[1] : (p == 1 ? mu * F1[1] : s12),
                     p == 2 ? mu * F1[2] : s12};
    const T r2[3] = {p == 0 ? mu * F2[0] : s12, d2 == 1 ? ltr + m2 * F2[1] : (p == 1 ? mu * F2[1] : s12),
                     d2 == 2 ? ltr + m2 * F2[2] : (p == 2 ? mu * F2[2] : s12)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T u1 = K.w1[d1] * r1[c], u2 = K.w1[d2] * r2[c];
      y[c][n1] = p < d1 ? u1 : y[c][n1] + u1;
      y[c][n2] = p < d2 ? u2 : y[c][n2] + u2;
    }
  }
  // trilinear modes: g[c][7] = 8/9 b[c][7] (kappa_c^2 (lam + 2 mu) + mu sum_{d != c} kappa_d^2)
  {
    const T lpm = lam + mu, t = mu * K.kks;
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c][7] = (K.kk[c] * lpm + t) * b[c][7];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) y[c][0] = T(0);
  if (MASS && !lumped) {
    // consistent mass: g[c][m] += mc / (8 3^|m|) b[c][m]
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int w = (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1);
      const T f = mc * T(w == 0 ? 1.0 / 8.0 : w == 1 ? 1.0 / 24.0 : w == 2 ? 1.0 / 72.0 : 1.0 / 216.0);
#pragma unroll
      for (int c = 0; c < 3; ++c) y[c][m] += f * b[c][m];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) hadamard8_transposed(y[c]);
  if (MASS && lumped) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < 8; ++q) y[c][q] += mc * xs[c * NODES + ln[q]];
  }
}

// SSPbrick on an axis-aligned box (ssp_brick.hpp; element_operator_apply with its element term,
// element_operator.hpp:82-130): the one-point mean strain is the gradient of the linear modes, so
// the mean-strain stiffness is box_constant_strain exactly; the stabilisation Kstab = Mben^T
// interior Mben acts on the hourglass modes only, whose vectors on a box are the modes xi eta,
// eta zeta, xi zeta, xi eta zeta (3, 6, 5, 7 here): Mben v = (H v)_m / 8, so the twelve modal
// values go through one symmetric 12 x 12 product with S = interior / 64 (per parameter set,
// packed upper triangle), scaled by the operator factor fE. The lumped mass mc x at each node.
template <typename T, bool MASS, int NODES, typename SPtr>
__device__ __forceinline__ void box_apply_ssp(const T* xs, const int (&ln)[8], T lam, T mu, T mc, T fE, SPtr S,
                                              const BoxConst<T>& K, T (&y)[3][8]) {
  T b[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int q = 0; q < 8; ++q) b[c][q] = xs[c * NODES + ln[q]];
    hadamard8(b[c]);
  }
  box_constant_strain(b, lam, mu, K, y);
  constexpr int kHourglass[4] = {3, 6, 5, 7};
  T w[12], z[12];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      w[3 * g + c] = b[c][kHourglass[g]];
      z[3 * g + c] = T(0);
    }
  int idx = 0;
#pragma unroll
  for (int r = 0; r < 12; ++r)
#pragma unroll
    for (int q = r; q < 12; ++q) {
      const T sv = static_cast<T>(S[idx++]);
      z[r] += sv * w[q];
      if (q != r) z[q] += sv * w[r];
    }
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int c = 0; c < 3; ++c) y[c][kHourglass[g]] = fE * z[3 * g + c];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    y[c][0] = T(0);
    hadamard8_transposed(y[c]);
  }
  if (MASS) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < 8; ++q) y[c][q] += mc * xs[c * NODES + ln[q]];
  }
}

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) {
  return __shfl_up(v, static_cast<unsigned>(d), 64);
}

// The value of lane t - 1 by a DPP row shift (row_shr:1). Lane 0 of each 16-lane row receives
// 0 (bound_ctrl); the callers weigh the result by 0 where the cell has no x-neighbour (cx = 0:
// lanes 0 and 8 of every row).
__device__ __forceinline__ double lane_up1(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x111, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x111, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_up1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x111, 0xf, 0xf, true));
}

constexpr int kPlane = (kTileX + 1) * (kTileY + 1);  // nodes of one z-plane of a tile
constexpr int kTileThreads = kTileX * kTileY * 4;     // one wave per z-plane of cells, four planes a pass

// Tiles of 8 x 8 x TZ cells (TZ = 4 or 8): 256 threads, TZ / 4 passes of four cell planes (a
// thread takes cell (cx, cy, wave + 4 pass)). Every tile owns full blocks of kNodes words and
// slots and kCells cell parameters (full-tile order, x + 8 (y + 8 z)).
template <int TZ>
struct TileGeom {
  static constexpr int kNodes = kPlane * (TZ + 1);
  static constexpr int kCells = kTileX * kTileY * TZ;
  static constexpr int kPasses = TZ / 4;
  static constexpr int kRounds = (kNodes + kTileThreads - 1) / kTileThreads;
  // top-face sums of the cell planes kept in LDS: a ring of kRing planes. With two passes, pass
  // 1 writes the slots of planes 0-2 only where pass 0 read them in the same wave, and plane
  // 3's slot (read by pass 1) is not written again.
  static constexpr int kRing = TZ == 4 ? 4 : 5;
};

// What one thread fetches of a tile before its cells are computed: the words of tile nodes
// t + 256 r (the partial-sum slot in place of the first equation for a shared node), x at their
// free DOFs, and its cells' parameter sets.
template <typename T, int TZ>
struct TileFetch {
  uint32_t w[TileGeom<TZ>::kRounds];
  T v[TileGeom<TZ>::kRounds][3];
  int pidx[TileGeom<TZ>::kPasses];
};

// A strictly-interior node (x, y, z in 1..7, 1..7, 1..TZ-1) of a regular tile: its word from the
// tile's interior base (three free DOFs, interior order x fastest), or 0xffffffff.
template <int TZ>
__device__ __forceinline__ uint32_t interior_word(int ib, int ln) {
  const int x = ln % (kTileX + 1), y = (ln / (kTileX + 1)) % (kTileY + 1), z = ln / kPlane;
  if (ib < 0 || x < 1 || x >= kTileX || y < 1 || y >= kTileY || z < 1 || z >= TZ) return 0xffffffffu;
  return (7u << 28) | static_cast<uint32_t>(ib + 3 * ((x - 1) + (kTileX - 1) * ((y - 1) + (kTileY - 1) * (z - 1))));
}

template <typename T, bool DIAG, int TZ>
__device__ __forceinline__ void fetch_tile(const BoxArgs<T>& a, int tile, int t, TileFetch<T, TZ>& f) {
  using G = TileGeom<TZ>;
  const uint32_t* words = a.words + static_cast<size_t>(tile) * G::kNodes;
  const int32_t* slots = a.slots + static_cast<size_t>(tile) * G::kNodes;
  const int ib = a.use_interior ? __builtin_amdgcn_readfirstlane(a.tiles[tile].interior_base) : -1;
  uint32_t w[G::kRounds];
  int32_t s[G::kRounds];
#pragma unroll
  for (int r = 0; r < G::kRounds; ++r) {
    const int ln = t + r * kTileThreads;
    const uint32_t iw = interior_word<TZ>(ib, ln);
    const bool load = ln < G::kNodes && iw == 0xffffffffu;
    w[r] = load ? words[ln] : ln < G::kNodes ? iw : 0u;
    s[r] = load ? slots[ln] : 0;
  }
#pragma unroll
  for (int p = 0; p < G::kPasses; ++p) f.pidx[p] = a.cell_param[static_cast<size_t>(tile) * G::kCells + t + kTileThreads * p];
#pragma unroll
  for (int r = 0; r < G::kRounds; ++r) {
    const uint32_t mask = DIAG ? 0u : (w[r] >> 28) & 7u;  // 0 past the tile's nodes
    const T* src = a.x + (w[r] & 0x0fffffffu);
    if (mask == 7u) {  // a node with three free DOFs (almost all): three consecutive equations
#pragma unroll
      for (int d = 0; d < 3; ++d) f.v[r][d] = src[d];
    } else {
      int j = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        f.v[r][d] = (mask & (1u << d)) ? src[j] : T(0);
        j += (mask >> d) & 1u;
      }
    }
    f.w[r] = (w[r] & 0x80000000u) ? ((w[r] & 0xf0000000u) | static_cast<uint32_t>(s[r])) : w[r];
  }
}

struct NoHook {
  __device__ void operator()() const {}
};

// One pass of a tile: cell (cx, cy, cz = wave + 4 pass). MODE: 0 apply without mass terms, 1
// apply with mass terms, 2 diagonal; 3, 4, 5 the same for SSPbrick cells; 6, 7, 8 the same on a
// rectilinear lattice (each cell's spacings from the tile's origin). The cell's results are
// summed per node and committed.
//
// Every lane runs the full 8 x 8 plane: a cell past the end of a cut tile has no parameter set
// (zero results) and a node past it has no word (nothing fetched, nothing committed), so the
// node sums need no bounds. Node (x, y, z) of the tile sits at x + 9 (y + 9 z') in LDS, with
// z' = z (RING = 0: every node plane of the tile held) or z mod RING (a ring of RING planes).
// mid() runs after the barrier that ends the pass's reads of x, before its commits.
template <typename T, int MODE, int TZ, int RING = 0, typename Mid = NoHook>
__device__ __forceinline__ void box_tile(const BoxArgs<T>& a, int pidx, const T* xs, const uint32_t* wsh, T* top,
                                         int t, int pass, int tile, const Mid& mid = Mid{}) {
  using G = TileGeom<TZ>;
  constexpr bool DIAG = MODE % 3 == 2;
  constexpr bool SSP = MODE >= 3 && MODE < 6;
  constexpr bool RECT = MODE >= 6;
  constexpr int NX = kTileX + 1;
  constexpr int LDS_NODES = RING ? RING * kPlane : G::kNodes;  // component stride of xs
  auto lp = [](int z) { return RING ? z % RING : z; };
  const int cx = t % kTileX, cy = (t / kTileX) % kTileY;
  const int cz = __builtin_amdgcn_readfirstlane(t / (kTileX * kTileY)) + 4 * pass;  // one plane per wave
  T out[3][8];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int q = 0; q < 8; ++q) out[c][q] = T(0);
  if (pidx != 0xffff) {
    const DevParam& p = a.params[pidx];
    double fD, cM;
    operator_factors(p, a.k, fD, cM);
    // the cell's geometry: the lattice's cell size, or on a rectilinear lattice the spacings of
    // the cell's plane intervals
    double hc[3] = {a.hx, a.hy, a.hz};
    BoxConst<T> Kc = a.K;
    if constexpr (RECT) {
      const TileHeader& th = a.tiles[tile];
      const double* kh[3] = {a.kx + 2 * (th.ox + cx), a.ky + 2 * (th.oy + cy), a.kz + 2 * (th.oz + cz)};
      double ks = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double kap = kh[d][0];
        hc[d] = kh[d][1];
        Kc.kap[d] = static_cast<T>(kap);
        Kc.w0[d] = static_cast<T>(8.0 * kap);
        Kc.w1[d] = static_cast<T>(8.0 / 3.0 * kap);
        Kc.kk[d] = static_cast<T>(8.0 / 9.0 * kap * kap);
        ks += kap * kap;
      }
      Kc.kks = static_cast<T>(8.0 / 9.0 * ks);
    }
    const double dvol = hc[0] * hc[1] * hc[2] * 0.125;
    if constexpr (DIAG && SSP) {
      // diag(K_e) per corner, kept per parameter set, and the lumped mass
      const double* dk = a.ssp + static_cast<size_t>(pidx) * kSspBox + 78;
      const double mass = cM * p.rho * dvol;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int q = 0; q < 8; ++q) out[c][q] = static_cast<T>(fD * dk[8 * c + q] + mass);
    } else if constexpr (DIAG) {
      // diag(K_e): Gauss sums of (dN/dx_d)^2 are 8 / (9 h_d^2) at every node of a box; of N^2, 8/27
      const double S[3] = {8.0 / (9.0 * hc[0] * hc[0]), 8.0 / (9.0 * hc[1] * hc[1]), 8.0 / (9.0 * hc[2] * hc[2])};
      const double lam = fD * p.lam, mu = fD * p.mu;
      const double mass = cM * p.rho * dvol * (p.mass_type == 0 ? 8.0 / 27.0 : 1.0);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double dc = dvol * ((lam + 2.0 * mu) * S[c] + mu * (S[0] + S[1] + S[2] - S[c])) + mass;
#pragma unroll
        for (int q = 0; q < 8; ++q) out[c][q] = static_cast<T>(dc);
      }
    } else {
      int ln[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) ln[q] = (cx + (q & 1)) + NX * (cy + ((q >> 1) & 1)) + kPlane * lp(cz + (q >> 2));
      if constexpr (SSP) {
        const T lam = static_cast<T>(fD * p.lam * dvol), mu = static_cast<T>(fD * p.mu * dvol);
        const T mc = static_cast<T>(cM * p.rho * dvol), fE = static_cast<T>(fD);
        // one parameter set across the wave (the usual case): its constants by scalar loads
        const int pu = __builtin_amdgcn_readfirstlane(pidx);
        if (__all(pidx == pu)) {
          using ConstPtr = const __attribute__((address_space(4))) double*;
          box_apply_ssp<T, MODE == 4, LDS_NODES>(xs, ln, lam, mu, mc, fE,
                                                 (ConstPtr)(a.ssp + static_cast<size_t>(pu) * kSspBox), a.K, out);
        } else {
          box_apply_ssp<T, MODE == 4, LDS_NODES>(xs, ln, lam, mu, mc, fE, a.ssp + static_cast<size_t>(pidx) * kSspBox,
                                                 a.K, out);
        }
      } else
        box_apply<T, MODE % 3 == 1, LDS_NODES>(xs, ln, static_cast<T>(fD * p.lam * dvol),
                                               static_cast<T>(fD * p.mu * dvol), static_cast<T>(cM * p.rho * dvol),
                                               p.mass_type != 0, Kc, out);
    }
  }
  // in-plane sums, x first then y: the cell's corners a (cx, cy), b (cx+1, cy), c (cx, cy+1),
  // d (cx+1, cy+1) of its bottom (k = 0) or top (k = 1) face. Row sums p = a + b(t-1) at node
  // (cx, cy) and q = c + d(t-1) at (cx, cy+1); then node (cx, cy) gets m = p + q(t-8), node
  // (cx+1, cy) e = b + d(t-8). A missing neighbour (cx = 0 or cy = 0) enters with factor 0 (a
  // fused multiply-add by 1 is the plain sum). The last column and row of cells also keep e, q
  // and d for the x = 8 column, y = 8 row and their corner. Top-face sums go to the LDS ring
  // (the plane above reads them), bottom-face sums stay in registers.
  const T fx = cx > 0 ? T(1) : T(0), fy = cy > 0 ? T(1) : T(0);
  const bool xe = cx == kTileX - 1, ye = cy == kTileY - 1;
  const int pi[4] = {cx + NX * cy, (cx + 1) + NX * cy, cx + NX * (cy + 1), (cx + 1) + NX * (cy + 1)};
  const bool has[4] = {true, xe, ye, xe && ye};
  auto slot = [&](int c, int plane) { return (c * G::kRing + plane % G::kRing) * kPlane; };
  T F0[3][4];  // bottom face: [component][main, x edge, y edge, corner]
#pragma unroll
  for (int k = 1; k >= 0; --k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T oa = out[c][4 * k], ob = out[c][1 + 4 * k], oc = out[c][2 + 4 * k], od = out[c][3 + 4 * k];
      const T p = fma_t(lane_up1(ob), fx, oa);
      const T q = fma_t(lane_up1(od), fx, oc);
      const T m = fma_t(shfl_up(q, kTileX), fy, p);
      const T e = fma_t(shfl_up(od, kTileX), fy, ob);
      const T f[4] = {m, e, q, od};
      if (k == 1) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (has[r]) top[slot(c, cz) + pi[r]] = f[r];
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) F0[c][r] = f[r];
      }
    }
  __syncthreads();
  mid();
  // commit: plane z = cz (bottom face + the plane below's top face), and for the last plane of
  // cells also plane z = TZ (its own top face)
  auto commit = [&](int ln, const T (&val)[3]) {
    const uint32_t w = wsh[ln];
    const uint32_t idx = w & 0x0fffffffu;
    const bool shared = (w & 0x80000000u) != 0u;
    const uint32_t mask = shared ? 7u : (w >> 28) & 7u;
    // a shared node: this tile's partial-sum slot (always three values, never accumulated)
    T* dst = shared ? a.partial + static_cast<size_t>(idx) * 3 : a.y + idx;
    if (mask == 7u) {
      T v[3] = {val[0], val[1], val[2]};
      if (a.accumulate && !shared)
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] = dst[d] + v[d];
#pragma unroll
      for (int d = 0; d < 3; ++d) dst[d] = v[d];
    } else if (mask != 0u) {
      int j = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (!(mask & (1u << d))) continue;
        T* o = dst + j++;
        *o = a.accumulate ? *o + val[d] : val[d];
      }
    }
  };
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    if (!has[r]) continue;
    T val[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) val[c] = cz > 0 ? F0[c][r] + top[slot(c, cz - 1) + pi[r]] : F0[c][r];
    commit(pi[r] + kPlane * lp(cz), val);
    if (cz == TZ - 1) {
      T up[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) up[c] = top[slot(c, cz) + pi[r]];
      commit(pi[r] + kPlane * lp(cz + 1), up);
    }
  }
}

// One workgroup per tile (PERSIST: each workgroup walks tiles blockIdx.x, + gridDim.x, ...
// and fetches the next tile's words and x while it computes the current one). WAVES: the
// minimum waves per SIMD the register allocation must allow (0: the compiler's choice).
template <typename T, int MODE, int WAVES, bool PERSIST, int TZ>
__global__ __launch_bounds__(kTileThreads, WAVES > 0 ? WAVES : 1) void hex8_box_kernel(BoxArgs<T> a) {
  using G = TileGeom<TZ>;
  constexpr bool DIAG = MODE % 3 == 2;
  __shared__ T xs[3 * G::kNodes];               // x at the tile's nodes
  __shared__ uint32_t wsh[G::kNodes];           // node words (slot for shared nodes)
  __shared__ T top[3 * G::kRing * kPlane];      // top-face sums of the cell planes
  const int t = threadIdx.x;
  if (a.done && *a.done) return;
  int tile = blockIdx.x;
  TileFetch<T, TZ> f;
  fetch_tile<T, DIAG, TZ>(a, tile, t, f);
  for (;;) {
#pragma unroll
    for (int r = 0; r < G::kRounds; ++r) {
      const int ln = t + r * kTileThreads;
      if (ln < G::kNodes) {
        wsh[ln] = f.w[r];
#pragma unroll
        for (int d = 0; d < 3; ++d) xs[d * G::kNodes + ln] = f.v[r][d];
      }
    }
    __syncthreads();
    const int next = PERSIST ? tile + static_cast<int>(gridDim.x) : a.ntiles;
    TileFetch<T, TZ> fn = f;
    if (PERSIST && next < a.ntiles) fetch_tile<T, DIAG, TZ>(a, next, t, fn);
#pragma unroll
    for (int pass = 0; pass < G::kPasses; ++pass) box_tile<T, MODE, TZ>(a, f.pidx[pass], xs, wsh, top, t, pass, tile);
    if (next >= a.ntiles) break;
    __syncthreads();  // the LDS takes the next tile
    tile = next;
    f = fn;
  }
}

// Streamed tiles: the x and the words of only five node planes sit in LDS, in rings (node plane
// z at ring plane z mod 5, ring index = tile node index mod 405). The first five planes are
// fetched before the first pass; each later pass's four new planes are loaded while the pass
// before it commits, and stored over the planes that pass has finished with. 21 KB of LDS per
// workgroup against 30 KB with every plane held: seven resident workgroups per CU, not five.
constexpr int kRingPlanes = 5;
constexpr int kRingNodes = kRingPlanes * kPlane;

template <typename T>
struct PlaneFetch {
  uint32_t w[2];
  T v[2][3];
};

// tile nodes [n0, n0 + cnt) (whole node planes, cnt <= 512) into registers: thread t takes
// nodes n0 + t and n0 + t + 256
template <typename T, bool DIAG>
__device__ __forceinline__ void fetch_planes(const BoxArgs<T>& a, const uint32_t* words, const int32_t* slots, int n0,
                                             int cnt, int t, PlaneFetch<T>& f) {
  uint32_t w[2];
  int32_t s[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int j = t + r * kTileThreads;
    w[r] = j < cnt ? words[n0 + j] : 0u;
    s[r] = j < cnt ? slots[n0 + j] : 0;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const uint32_t mask = DIAG ? 0u : (w[r] >> 28) & 7u;
    const T* src = a.x + (w[r] & 0x0fffffffu);
    if (mask == 7u) {
#pragma unroll
      for (int d = 0; d < 3; ++d) f.v[r][d] = src[d];
    } else {
      int j = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        f.v[r][d] = (mask & (1u << d)) ? src[j] : T(0);
        j += (mask >> d) & 1u;
      }
    }
    f.w[r] = (w[r] & 0x80000000u) ? ((w[r] & 0xf0000000u) | static_cast<uint32_t>(s[r])) : w[r];
  }
}

template <typename T>
__device__ __forceinline__ void store_planes(const PlaneFetch<T>& f, int n0, int cnt, int t, T* xs, uint32_t* wsh) {
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int j = t + r * kTileThreads;
    if (j < cnt) {
      const int ln = (n0 + j) % kRingNodes;
      wsh[ln] = f.w[r];
#pragma unroll
      for (int d = 0; d < 3; ++d) xs[d * kRingNodes + ln] = f.v[r][d];
    }
  }
}

template <typename T, int MODE, int WAVES, int TZ>
__global__ __launch_bounds__(kTileThreads, WAVES > 0 ? WAVES : 1) void hex8_stream_kernel(BoxArgs<T> a) {
  using G = TileGeom<TZ>;
  constexpr bool DIAG = MODE == 2;
  static_assert(G::kPasses >= 2, "streamed tiles have at least two passes");
  __shared__ T xs[3 * kRingNodes];
  __shared__ uint32_t wsh[kRingNodes];
  __shared__ T top[3 * G::kRing * kPlane];
  const int t = threadIdx.x;
  if (a.done && *a.done) return;
  const int tile = blockIdx.x;
  const uint32_t* words = a.words + static_cast<size_t>(tile) * G::kNodes;
  const int32_t* slots = a.slots + static_cast<size_t>(tile) * G::kNodes;
  int pidx[G::kPasses];
#pragma unroll
  for (int p = 0; p < G::kPasses; ++p) pidx[p] = a.cell_param[static_cast<size_t>(tile) * G::kCells + t + kTileThreads * p];
  {
    PlaneFetch<T> f;
    fetch_planes<T, DIAG>(a, words, slots, 0, kRingNodes, t, f);
    store_planes(f, 0, kRingNodes, t, xs, wsh);
  }
  __syncthreads();
#pragma unroll
  for (int pass = 0; pass < G::kPasses; ++pass) {
    if (pass + 1 < G::kPasses) {
      // planes 4 pass + 5 .. 4 pass + 8 replace planes 4 pass .. 4 pass + 3 (ring slots)
      const int n0 = (4 * pass + kRingPlanes) * kPlane;
      PlaneFetch<T> f;
      box_tile<T, MODE, TZ, kRingPlanes>(a, pidx[pass], xs, wsh, top, t, pass, tile,
                                         [&] { fetch_planes<T, DIAG>(a, words, slots, n0, 4 * kPlane, t, f); });
      __syncthreads();  // this pass's commits have read the words of its planes
      store_planes(f, n0, 4 * kPlane, t, xs, wsh);
      __syncthreads();
    } else {
      box_tile<T, MODE, TZ, kRingPlanes>(a, pidx[pass], xs, wsh, top, t, pass, tile);
    }
  }
}

// y at the nodes several tiles touch: their tiles' partial sums added in colour order. One
// thread per shared node; a group of nodes with the same tiles keeps one block of slots per
// tile, so the threads of a wave read each block contiguously. The shared nodes' equations are
// consecutive in the internal numbering, in the same order.
template <typename T>
__global__ __launch_bounds__(256) void hex8_shared_sum_kernel(const uint32_t* __restrict__ word,
                                                              const int32_t* __restrict__ start,
                                                              const uint32_t* __restrict__ meta,
                                                              const T* __restrict__ partial, T* __restrict__ y,
                                                              int n, int accumulate, const int32_t* done) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (done && *done) return;
  const uint32_t w = word[i];
  const uint32_t mask = (w >> 28) & 7u;
  const int base = static_cast<int>(w & 0x0fffffffu);
  const int s0 = start[i];
  const uint32_t mt = meta[i];
  const int stride = static_cast<int>(mt & 0xffffffu), k = static_cast<int>(mt >> 24);
  T acc[3];
  {
    const T* p = partial + static_cast<size_t>(s0) * 3;
    int j = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      acc[d] = p[d];
      if (accumulate && (mask & (1u << d))) acc[d] = y[base + j] + acc[d];
      j += (mask >> d) & 1u;
    }
  }
  // the other tiles' partials (at most 8 tiles touch a node): loads first, then the adds in
  // colour order
  T v[7][3];
#pragma unroll
  for (int t = 1; t < 8; ++t)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      v[t - 1][d] = t < k ? partial[(static_cast<size_t>(s0) + static_cast<size_t>(t) * stride) * 3 + d] : T(0);
#pragma unroll
  for (int t = 1; t < 8; ++t)
    if (t < k)
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[d] += v[t - 1][d];
  int j = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (mask & (1u << d)) y[base + j++] = acc[d];
}

// y at tied equation sets (nodes that hold the same equations: the laminar ties of a soil
// column): the partial sums of every tile at every node of the set, one wave per set, the lanes
// striding over the set's slots, then a butterfly over the lanes (a fixed order).
template <typename T>
__global__ __launch_bounds__(256) void hex8_tied_sum_kernel(const uint32_t* __restrict__ word,
                                                            const int32_t* __restrict__ start,
                                                            const T* __restrict__ partial, T* __restrict__ y, int n,
                                                            int accumulate, const int32_t* done) {
  const int u = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x) / 64;  // one set per wave
  const int lane = threadIdx.x % 64;
  if (u >= n) return;  // the whole wave
  if (done && *done) return;
  const int s0 = start[u], s1 = start[u + 1];
  T acc[3] = {T(0), T(0), T(0)};
  for (int s = s0 + lane; s < s1; s += 64)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] += partial[static_cast<size_t>(s) * 3 + d];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] += __shfl_xor(acc[d], off, 64);
  if (lane != 0) return;
  const uint32_t w = word[u];
  const uint32_t mask = (w >> 28) & 7u;
  const int base = static_cast<int>(w & 0x0fffffffu);
  int j = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!(mask & (1u << d))) continue;
    T* dst = y + base + j++;
    *dst = accumulate ? *dst + acc[d] : acc[d];
  }
}

namespace {
// Workgroups of a tile kernel resident at once on the device (blocks per CU x CUs), cached.
int resident_tiles(fx_ctx* ctx, const void* kernel) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, int> cache;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(ctx->device, kernel);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTileThreads, 0) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return cache[key] = per_cu * ncu;
}
}  // namespace

template <typename T, bool DIAG>
int launch_box(fx_batch* b, const fx_coeffs& k, const T* x, T* y, int accumulate, const int32_t* done, int ref) {
  if (ref && !b->words_ref) {
    b->ctx->error = "operator: no reference-numbering words for this batch";
    return FX_EINVAL;
  }
  // the mass-free kernel unless some parameter set has a mass factor cM' = cM + cC alphaM
  bool mass = false;
  for (const DevParam& p : b->host_params) {
    double fD, cM;
    operator_factors(p, k, fD, cM);
    mass |= cM != 0.0;
  }
  // variant bits 0-1: the register budget of the mass-free kernel (the compiler's choice, or at
  // least 5, 7, 8 waves per SIMD: 84, 72, 64 VGPRs in FP64, the last two with a few spills; the
  // mass and diagonal kernels keep the compiler's choice); bit 2: persistent workgroups with
  // prefetch
  using Kernel = void (*)(BoxArgs<T>);
#define FX_BOX_TABLE(TZ)                                                                                         \
  {{hex8_box_kernel<T, 0, 0, false, TZ>, hex8_box_kernel<T, 0, 5, false, TZ>, hex8_box_kernel<T, 0, 7, false, TZ>, \
    hex8_box_kernel<T, 0, 8, false, TZ>, hex8_box_kernel<T, 0, 0, true, TZ>, hex8_box_kernel<T, 0, 5, true, TZ>,    \
    hex8_box_kernel<T, 0, 7, true, TZ>, hex8_box_kernel<T, 0, 8, true, TZ>},                                       \
   {hex8_box_kernel<T, 1, 0, false, TZ>, hex8_box_kernel<T, 1, 0, false, TZ>, hex8_box_kernel<T, 1, 0, false, TZ>, \
    hex8_box_kernel<T, 1, 0, false, TZ>, hex8_box_kernel<T, 1, 0, true, TZ>, hex8_box_kernel<T, 1, 0, true, TZ>,    \
    hex8_box_kernel<T, 1, 0, true, TZ>, hex8_box_kernel<T, 1, 0, true, TZ>},                                       \
   {hex8_box_kernel<T, 2, 0, false, TZ>, hex8_box_kernel<T, 2, 0, false, TZ>, hex8_box_kernel<T, 2, 0, false, TZ>, \
    hex8_box_kernel<T, 2, 0, false, TZ>, hex8_box_kernel<T, 2, 0, true, TZ>, hex8_box_kernel<T, 2, 0, true, TZ>,    \
    hex8_box_kernel<T, 2, 0, true, TZ>, hex8_box_kernel<T, 2, 0, true, TZ>}}
  static const Kernel table4[3][8] = FX_BOX_TABLE(4);
  static const Kernel table8[3][8] = FX_BOX_TABLE(8);
#undef FX_BOX_TABLE
  // bit 3: streamed tiles (rings of five node planes in LDS; 8-deep tiles). 16-deep tiles are
  // always streamed (bits 0-1 the register budget as above, bit 2 ignored).
#define FX_STREAM_TABLE(TZ)                                                                                         \
  {{hex8_stream_kernel<T, 0, 0, TZ>, hex8_stream_kernel<T, 0, 5, TZ>, hex8_stream_kernel<T, 0, 7, TZ>,             \
    hex8_stream_kernel<T, 0, 8, TZ>},                                                                              \
   {hex8_stream_kernel<T, 1, 0, TZ>, hex8_stream_kernel<T, 1, 0, TZ>, hex8_stream_kernel<T, 1, 0, TZ>,             \
    hex8_stream_kernel<T, 1, 0, TZ>},                                                                              \
   {hex8_stream_kernel<T, 2, 0, TZ>, hex8_stream_kernel<T, 2, 0, TZ>, hex8_stream_kernel<T, 2, 0, TZ>,             \
    hex8_stream_kernel<T, 2, 0, TZ>}}
  static const Kernel stream8[3][4] = FX_STREAM_TABLE(8);
  static const Kernel stream16[3][4] = FX_STREAM_TABLE(16);
#undef FX_STREAM_TABLE
  // SSPbrick lattices: one kernel per mode (4- or 8-deep tiles, the compiler's register budget)
  static const Kernel ssp4[3] = {hex8_box_kernel<T, 3, 0, false, 4>, hex8_box_kernel<T, 4, 0, false, 4>,
                                 hex8_box_kernel<T, 5, 0, false, 4>};
  static const Kernel ssp8[3] = {hex8_box_kernel<T, 3, 0, false, 8>, hex8_box_kernel<T, 4, 0, false, 8>,
                                 hex8_box_kernel<T, 5, 0, false, 8>};
  // rectilinear lattices: the same, each cell's spacings from its plane intervals
  static const Kernel rect4[3] = {hex8_box_kernel<T, 6, 0, false, 4>, hex8_box_kernel<T, 7, 0, false, 4>,
                                  hex8_box_kernel<T, 8, 0, false, 4>};
  static const Kernel rect8[3] = {hex8_box_kernel<T, 6, 0, false, 8>, hex8_box_kernel<T, 7, 0, false, 8>,
                                  hex8_box_kernel<T, 8, 0, false, 8>};
  const bool one = b->ssp || b->rect;  // a single kernel per mode
  const int m = DIAG ? 2 : mass ? 1 : 0;
  const int v = one ? 0 : (m == 0 ? b->box_variant : b->box_variant_mass) & 15;
  const bool stream = !one && (b->tile_z == 16 || ((v & 8) && b->tile_z == 8));
  const Kernel kernel = b->ssp            ? (b->tile_z == 8 ? ssp8 : ssp4)[m]
                        : b->rect         ? (b->tile_z == 8 ? rect8 : rect4)[m]
                        : b->tile_z == 16 ? stream16[m][v & 3]
                        : stream          ? stream8[m][v & 3]
                                          : (b->tile_z == 8 ? table8 : table4)[m][v & 7];
  int grid = b->ntiles;
  if ((v & 4) && !stream) {
    const int resident = resident_tiles(b->ctx, reinterpret_cast<const void*>(kernel));
    if (resident > 0 && resident < grid) grid = resident >= 8 ? resident / 8 * 8 : resident;
  }
  BoxArgs<T> a{};
  a.tiles = b->tiles;
  a.words = ref ? b->words_ref : b->words;
  a.slots = b->tile_slots;
  a.use_interior = ref ? 0 : 1;
  a.cell_param = b->cell_param;
  a.params = b->params;
  a.k = k;
  a.hx = b->hx;
  a.hy = b->hy;
  a.hz = b->hz;
  const double kap[3] = {0.25 / b->hx, 0.25 / b->hy, 0.25 / b->hz};
  for (int d = 0; d < 3; ++d) {
    a.K.kap[d] = static_cast<T>(kap[d]);
    a.K.w0[d] = static_cast<T>(8.0 * kap[d]);
    a.K.w1[d] = static_cast<T>(8.0 / 3.0 * kap[d]);
    a.K.kk[d] = static_cast<T>(8.0 / 9.0 * kap[d] * kap[d]);
  }
  a.K.kks = static_cast<T>(8.0 / 9.0 * (kap[0] * kap[0] + kap[1] * kap[1] + kap[2] * kap[2]));
  a.x = x;
  a.y = y;
  if constexpr (sizeof(T) == 8) {
    a.partial = b->partial64;
  } else {
    a.partial = b->partial32;
  }
  a.ntiles = b->ntiles;
  a.accumulate = accumulate;
  a.done = done;
  a.ssp = b->ssp_box;
  a.kx = b->kx;
  a.ky = b->ky;
  a.kz = b->kz;
  if (!a.partial) {
    b->ctx->error = "operator: no partial-sum buffer for this precision";
    return FX_EINVAL;
  }
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kTileThreads), 0, b->ctx->stream, a);
  if (b->nshared > 0)
    hipLaunchKernelGGL(hex8_shared_sum_kernel<T>, dim3((b->nshared + 255) / 256), dim3(256), 0, b->ctx->stream,
                       ref ? b->shared_word_ref : b->shared_word, b->shared_start, b->shared_meta, a.partial, y,
                       b->nshared, accumulate, done);
  if (b->ntied > 0)
    hipLaunchKernelGGL(hex8_tied_sum_kernel<T>, dim3((b->ntied + 3) / 4), dim3(256), 0, b->ctx->stream,
                       ref ? b->tied_word_ref : b->tied_word, b->tied_start, a.partial, y, b->ntied, accumulate, done);
  FX_HIP(b->ctx, hipGetLastError());
  return FX_OK;
}

template int launch_box<double, false>(fx_batch*, const fx_coeffs&, const double*, double*, int, const int32_t*, int);
template int launch_box<double, true>(fx_batch*, const fx_coeffs&, const double*, double*, int, const int32_t*, int);
template int launch_box<float, false>(fx_batch*, const fx_coeffs&, const float*, float*, int, const int32_t*, int);

}  // namespace fx
